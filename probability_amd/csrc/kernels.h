// Sampling kernels: one wavefront = one chain.
//
// * nuts_kernel: NoUTurnSampler.one_step (tensorflow_probability/python/
//   mcmc/nuts.py:321-1010) as a per-chain state machine.  A chain's control
//   flow is wave-uniform, so a chain that stops (U-turn, divergence) leaves
//   its loop instead of idling under a mask until the slowest chain of the
//   batch is done, which is what the lockstep tf.while_loop does
//   (nuts.py:404-422, 756-759).  Outputs are identical because TFP freezes a
//   stopped chain's outputs and every draw here is keyed per chain.
// * hmc_kernel: MetropolisHastings(UncalibratedHamiltonianMonteCarlo)
//   (mcmc/hmc.py:661-875, mcmc/metropolis_hastings.py:160-254) fused with
//   the sample_chain loop over n_iters iterations.
//
// State element e of a chain lives in lane e % 64, register slot e / 64;
// the per-wave LDS block holds the NUTS U-turn memory and target scratch.
#pragma once

#include "../../include/hipmc.h"
#include "common.h"
#include "targets.h"

namespace hipmc {

// Workgroups are dispatched round-robin over the 8 XCDs (block b runs on
// XCD b % 8).  Renumbering so that consecutive logical blocks share an XCD
// keeps the chains of one problem -- which read the same data -- on one L2.
__device__ inline int xcd_block(int b, int n) {
  const int x = b & 7, per = n >> 3, rem = n & 7;
  return x * per + (x < rem ? x : rem) + (b >> 3);
}

struct ChainId {
  int chain;      // row in the launch
  int prob;       // problem index (table row)
  int local;      // chain index inside its problem (keys the draws)
  uint32_t key;   // problem word of the Philox counter
};

__device__ inline ChainId chain_of(const hipmc_run_args& a, int c) {
  ChainId id;
  id.chain = c;
  id.prob = a.chain_problem ? a.chain_problem[c] : 0;
  id.local = a.chain_local ? a.chain_local[c] : c;
  id.key = a.problems[id.prob].key;
  return id;
}

// Launch-level or per-problem sampling stream of a chain.
struct Stream {
  Seed base;
  int64_t iter0;
  int n_iters;
  bool fold;
  // trace
  void* trace;
  int64_t trace_start;
  int trace_every, n_trace;
  bool own_trace;  // dense per-problem trace [row][local][dim]
};

__device__ inline Stream stream_of(const hipmc_run_args& a, int prob) {
  Stream s;
  s.fold = a.fold_seed != 0;
  if (a.streams) {
    const hipmc_stream& st = a.streams[prob];
    s.base = Seed{st.seed[0], st.seed[1]};
    s.iter0 = st.iter0;
    s.n_iters = st.n_iters < a.n_iters ? st.n_iters : a.n_iters;
    s.trace = st.trace;
    s.trace_start = st.trace_start;
    s.trace_every = st.trace_every > 0 ? st.trace_every : 1;
    s.n_trace = st.n_trace;
    s.own_trace = true;
  } else {
    s.base = Seed{a.seed[0], a.seed[1]};
    s.iter0 = a.iter0;
    s.n_iters = a.n_iters;
    s.trace = a.trace_state;
    s.trace_start = a.trace_start;
    s.trace_every = a.trace_every > 0 ? a.trace_every : 1;
    s.n_trace = a.n_trace;
    s.own_trace = false;
  }
  return s;
}

__device__ inline Seed iter_seed(const Stream& s, int it) {
  return s.fold ? fold_in(s.base, (uint64_t)(s.iter0 + it)) : s.base;
}

// Trace row of launch iteration `it`, or -1 if it is not recorded.
__device__ inline int64_t trace_row(const Stream& s, int it) {
  if (!s.trace) return -1;
  const int64_t g = s.iter0 + it - s.trace_start;
  if (g < 0 || g % s.trace_every) return -1;
  const int64_t r = g / s.trace_every;
  return r < s.n_trace ? r : -1;
}

template <typename T, int K>
__device__ inline void store_row(T* row, const T (&v)[K], int dim);

template <typename T, int K>
__device__ inline void trace_store(const Stream& s, const hipmc_run_args& a,
                                   const ChainId& id, int dim, int it,
                                   const T (&v)[K]) {
  const int64_t tr = trace_row(s, it);
  if (tr < 0) return;
  if (s.own_trace) {
    const hipmc_problem& pb = a.problems[id.prob];
    store_row((T*)s.trace + ((size_t)tr * pb.n_chains + id.local) * dim, v,
              dim);
  } else {
    store_row((T*)s.trace + ((size_t)tr * a.n_chains + id.chain) *
                                a.dim_stride, v, dim);
  }
}

// Standard-normal momentum of the chain: element e uses Philox block e/4,
// word pair (e%4)/2, cos for even e, sin for odd e.
template <typename T, int K>
__device__ inline void draw_momentum(Seed s, const ChainId& id, int dim,
                                     T (&p)[K]) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = k * 64 + lane;
    const U4 w = philox(U4{(uint32_t)(e >> 2), (uint32_t)id.local, id.key,
                           kPurposeMomentum},
                        s.k0, s.k1);
    const bool hi_pair = (e & 2) != 0;
    const uint32_t wa = hi_pair ? w.z : w.x;
    const uint32_t wb = hi_pair ? w.w : w.y;
    double n0, n1;
    box_muller(wa, wb, n0, n1);
    const double n = (e & 1) ? n1 : n0;
    p[k] = (e < dim) ? (T)n : T(0);
  }
}

template <typename T, int K>
__device__ inline double half_sq_norm(const T (&p)[K]) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) s += (double)p[k] * (double)p[k];
  return 0.5 * wave_sum(s);
}

template <typename T, int K>
__device__ inline T dot(const T (&u)[K], const T (&v)[K]) {
  T s = T(0);
#pragma unroll
  for (int k = 0; k < K; ++k) s += u[k] * v[k];
  return wave_sum(s);
}

template <typename T, int K>
__device__ inline void copy(T (&d)[K], const T (&s)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) d[k] = s[k];
}

template <typename T, int K>
__device__ inline void load_row(T (&v)[K], const T* row, int dim) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = k * 64 + lane;
    v[k] = (e < dim) ? row[e] : T(0);
  }
}

template <typename T, int K>
__device__ inline void store_row(T* row, const T (&v)[K], int dim) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = k * 64 + lane;
    if (e < dim) row[e] = v[k];
  }
}

// `num_steps` leapfrog steps with the reference op order
// (leapfrog_integrator.py:280-309, 319-362), positions moving with the
// velocity prec * p of the momentum distribution (identity: prec = 1):
//   p += (0.5*eps)*g; repeat: q += eps*(prec*p); (t, g) = grad(q);
//   p += eps*g; p -= (0.5*eps)*g.
template <typename T, int K, int FAM, int G>
__device__ inline void leapfrog(Target<T, K, FAM, G>& tg, T eps, int num_steps,
                                const T (&m)[K], T (&q)[K], T (&p)[K],
                                T (&g)[K], double& t) {
  const T half = T(0.5) * eps;
#pragma unroll
  for (int k = 0; k < K; ++k) p[k] = p[k] + half * g[k];
  for (int s = 0; s < num_steps; ++s) {
#pragma unroll
    for (int k = 0; k < K; ++k) q[k] = q[k] + eps * (m[k] * p[k]);
    t = tg.eval(q, g);
#pragma unroll
    for (int k = 0; k < K; ++k) p[k] = p[k] + eps * g[k];
  }
#pragma unroll
  for (int k = 0; k < K; ++k) p[k] = p[k] - half * g[k];
}

// Diagonal momentum precision of the chain (hipmc_run_args
// .momentum_precision; identity when absent): prec[e] for the chain's
// elements (1 past dim), and the momentum density's normalising terms
// -0.5 dim log(2 pi) + 0.5 sum log prec, which the preconditioned NUTS
// includes in its energy (preconditioned_nuts.py:1091-1099, the log-density
// of MultivariateNormalPrecisionFactorLinearOperator).
template <typename T, int K>
__device__ inline bool load_precision(const hipmc_run_args& a,
                                      const ChainId& id, int dim, T (&m)[K],
                                      double& kconst) {
  const int lane = lane_id();
  kconst = 0.0;
  if (!a.momentum_precision) {
#pragma unroll
    for (int k = 0; k < K; ++k) m[k] = T(1);
    return false;
  }
  const T* row = (const T*)a.momentum_precision + (size_t)id.prob * a.dim_stride;
  double lsum = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int e = k * 64 + lane;
    m[k] = e < dim ? row[e] : T(1);
    lsum += e < dim ? log((double)m[k]) : 0.0;
  }
  kconst = 0.5 * wave_sum(lsum) - 0.5 * (double)dim * log(2.0 * M_PI);
  return true;
}

// Momentum of the preconditioned kernels: p = z / sqrt(prec)
// (MultivariateNormalPrecisionFactorLinearOperator samples x = F^-T w with
// F = diag(sqrt(prec)), mvn_precision_factor_linop.py:171-211).
template <typename T, int K>
__device__ inline void precondition_momentum(const T (&m)[K], T (&p)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) p[k] = p[k] / sqrt(m[k]);
}

// 0.5 * sum(prec * p^2), float64 (the kinetic energy without the normalising
// terms).
template <typename T, int K>
__device__ inline double half_kinetic(const T (&m)[K], const T (&p)[K]) {
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < K; ++k) s += (double)p[k] * (double)(m[k] * p[k]);
  return 0.5 * wave_sum(s);
}

// velocity prec * p
template <typename T, int K>
__device__ inline void velocity(T (&v)[K], const T (&m)[K], const T (&p)[K]) {
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = m[k] * p[k];
}

// Lane-wise copies of a register array to and from an LDS slot of K*64
// elements: each lane writes and reads back only its own elements, so no
// wave barrier is needed between them.
template <typename T, int K>
__device__ inline void park(T* slot, const T (&v)[K]) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < K; ++k) slot[k * 64 + lane] = v[k];
}
template <typename T, int K>
__device__ inline void unpark(T (&v)[K], const T* slot) {
  const int lane = lane_id();
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = slot[k * 64 + lane];
}

template <typename T>
__device__ inline T load_step(const hipmc_run_args& a, int c) {
  const int idx = a.chain_step ? a.chain_step[c] : 0;
  return ((const T*)a.step_size)[idx];
}

template <typename T>
__device__ inline void put(void* base, size_t i, T v) {
  if (base) ((T*)base)[i] = v;
}


// Waves sharing one chain: only GLM launches split a chain (at most 8,
// which at two waves per SIMD is one chain per CU).
template <int FAM>
__host__ __device__ inline int waves_per_chain(const hipmc_run_args& a) {
  if (FAM != kFamilyGlm || a.waves_per_chain <= 1) return 1;
  return a.waves_per_chain < 8 ? a.waves_per_chain : 8;
}

// ------------------------------------------------------------------ NUTS
// GLM chains hold their long-lived vectors in LDS ("parked"): the gradient
// evaluation keeps many loads in flight and needs most of the 256 registers
// that two waves per SIMD allow.  The trajectory ends, the candidate, the
// transition's start and the momentum sum change once per doubling and sit
// in LDS for the whole leaf loop (kTreeSlots slots, one copy per chain); the
// leaf's momentum, momentum sum and subtree candidate are parked around
// each gradient evaluation (kLeafSlots slots per wave).  The other families
// keep everything in registers (one wave per SIMD), except float32 states
// wider than 512 elements, which park the same way (at 17 slots and depth 10
// the U-turn memory, parked vectors and target scratch take 157 KiB of the
// workgroup's 160 KiB); float64 ones could not hold both in LDS.
constexpr int kTreeSlots = 11;
constexpr int kLeafSlots = 4;
template <typename T, int K, int FAM>
constexpr bool parks() {
  return FAM == kFamilyGlm || (K > 8 && sizeof(T) == 4);
}

// LDS of one chain (elements of T): U-turn memory velocity [max_depth][K*64]
// and cumulative momentum [max_depth][K*64], the parked tree vectors
// [kTreeSlots][K*64] (GLM); then per wave the target's scratch and the
// parked leaf vectors [kLeafSlots][K*64] (GLM).
template <typename T, int K, int FAM>
__host__ __device__ constexpr size_t nuts_chain_lds_elems(int max_depth) {
  return (size_t)2 * max_depth * K * 64 +
         (parks<T, K, FAM>() ? (size_t)kTreeSlots * K * 64 : 0);
}
template <typename T, int K, int FAM>
__host__ __device__ constexpr size_t nuts_wave_lds_elems() {
  return target_scratch_elems<T, K, FAM>() +
         (parks<T, K, FAM>() ? (size_t)kLeafSlots * K * 64 : 0);
}
// LDS of one wave when every wave holds its own chain.
template <typename T, int K, int FAM>
__host__ __device__ constexpr size_t nuts_lds_elems(int max_depth) {
  return nuts_chain_lds_elems<T, K, FAM>(max_depth) +
         nuts_wave_lds_elems<T, K, FAM>();
}

// One chain's transitions (the body of nuts_kernel).  BM: the target's
// bijector code path; U1: one leapfrog step per tree leaf
// (unrolled_leapfrog_steps == 1).  Specialising both keeps the leaf free of
// run-time branches and of the registers the other paths would hold.
// With a momentum precision (the preconditioned NUTS,
// preconditioned_nuts.py:330-1010) the U-turn memory holds velocities and
// the U-turn checks take momentum sums against velocities; with none, the
// velocity is the momentum (nuts.py:541-1010).
// W > 1: the W waves of one block share the chain (GLMs); they run the same
// control flow on identical values, and their parked tree vectors are one
// copy guarded by a block barrier before each write.
template <typename T, int K, int FAM, int BM, bool U1, int G = 1>
__device__ __attribute__((always_inline)) inline void nuts_chain(
    const hipmc_run_args& a, Target<T, K, FAM, G>& tg, const ChainId& id,
    int c, int wic, int W, T* memP, T* memC, T* tree, T* leafp,
    const hipmc_problem& pb) {
  constexpr bool PARK = parks<T, K, FAM>();
  constexpr size_t S = (size_t)K * 64;
  const int lane = lane_id();
  const int max_depth = a.max_tree_depth;
  const int dim = pb.dim;
  const int crow = a.chain_row ? a.chain_row[c] : c;
  const size_t row = (size_t)crow * a.dim_stride;
  const double max_de = a.max_energy_diff;
  const int nunroll = a.unrolled_leapfrog_steps;

  T m[K];
  double kconst;
  const bool pre = load_precision(a, id, dim, m, kconst);
  T q0[K], g0[K];
  load_row(q0, (const T*)a.state + row, dim);
  load_row(g0, (const T*)a.grad + row, dim);
  double t0 = (double)((const T*)a.tlp)[crow];
  const T eps = load_step<T>(a, c);

  const Stream stream = stream_of(a, id.prob);
  for (int it = 0; it < stream.n_iters; ++it) {
    const Seed sd = iter_seed(stream, it);
    T pL[K], qL[K], gL[K], pR[K], qR[K], gR[K];
    draw_momentum(sd, id, dim, pL);
    if (pre) precondition_momentum(m, pL);
    const double h0 = t0 - half_kinetic(m, pL);  // compute_hamiltonian, -H
    copy(qL, q0);
    copy(gL, g0);
    copy(pR, pL);
    copy(qR, q0);
    copy(gR, g0);
    double tL = t0, tR = t0;
    // multinomial candidate (nuts.py:361-366)
    T qC[K], gC[K];
    copy(qC, q0);
    copy(gC, g0);
    double tC = t0, eC = h0, wC = 0.0;
    T rho[K];
    copy(rho, pL);  // momentum_sum
    double eds = 0.0;
    int lf = 0;
    bool cont = true, notdiv = true, is_acc = false;
    int depth = 0;
    for (; depth < max_depth && cont; ++depth) {
      // _loop_tree_doubling (nuts.py:541-711)
      const U4 tw = philox(U4{0u, (uint32_t)id.local, id.key,
                              kPurposeTree | (uint32_t)depth},
                           sd.k0, sd.k1);
      const bool dir = (tw.x >> 31) != 0;
      const T tree_u = uniform_from<T>(tw.y, tw.z);
      const T e = dir ? eps : -eps;
      T q[K], p[K], g[K];
      double t;
      if (dir) {
        copy(q, qR); copy(p, pR); copy(g, gR); t = tR;
      } else {
        copy(q, qL); copy(p, pL); copy(g, gL); t = tL;
      }
      if (PARK) {
        if (W > 1) __syncthreads();  // every wave has read the last copy
        park(tree + 0 * S, pL); park(tree + 1 * S, qL); park(tree + 2 * S, gL);
        park(tree + 3 * S, pR); park(tree + 4 * S, qR); park(tree + 5 * S, gR);
        park(tree + 6 * S, qC); park(tree + 7 * S, gC); park(tree + 8 * S, q0);
        park(tree + 9 * S, g0); park(tree + 10 * S, rho);
      }
      // _build_sub_tree (nuts.py:713-791)
      T qS[K], gS[K];
      copy(qS, q);
      copy(gS, g);
      double tS = t, eS = t, wS = neg_inf();
      T cs[K];
#pragma unroll
      for (int k = 0; k < K; ++k) cs[k] = T(0);
      double eds_sub = 0.0;
      int lf_sub = 0;
      bool csub = true, nd = notdiv;
      const int nsteps = 1 << depth;
      // Per-lane buffers for 64 consecutive leaves: lane l holds leaf
      // (chunk + l)'s log1p(-U) (drawn for the whole chunk at once) and its
      // energy difference, whose exp(min(dH, 0)) terms are summed per chunk.
      double lu_buf = 0.0, de_buf = 0.0;
      bool de_ok = false;
      for (int i = 0; i < nsteps; ++i) {
        // _loop_build_sub_tree (nuts.py:793-946)
        const int li = i & 63;
        if (li == 0) {
          const U4 lw = philox(U4{(uint32_t)(i + lane), (uint32_t)id.local,
                                  id.key, kPurposeLeaf | (uint32_t)depth},
                               sd.k0, sd.k1);
          lu_buf = (double)log1m(uniform_from<T>(lw.x, lw.y));
        }
        double energy;
        T v[K];
        if (U1 || nunroll == 1) {
          const T half = T(0.5) * e;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            p[k] = p[k] + half * g[k];
            q[k] = q[k] + e * (m[k] * p[k]);
          }
          if (PARK) {
            park(leafp + 0 * S, p); park(leafp + 1 * S, cs);
            park(leafp + 2 * S, qS); park(leafp + 3 * S, gS);
          }
          double lpu;
          const double lpl = tg.template eval_partial<BM>(q, g, lpu);
          if (PARK) {
            unpark(p, leafp + 0 * S); unpark(cs, leafp + 1 * S);
            unpark(qS, leafp + 2 * S); unpark(gS, leafp + 3 * S);
          }
          double kin = 0.0;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            p[k] = p[k] + e * g[k];
            p[k] = p[k] - half * g[k];
            v[k] = m[k] * p[k];
            kin += (double)p[k] * (double)v[k];
          }
          double slp, skin;
          wave_sum2(lpl, kin, slp, skin);
          t = slp + lpu;
          energy = t - 0.5 * skin;
        } else {
          if (PARK) {
            park(leafp + 1 * S, cs); park(leafp + 2 * S, qS);
            park(leafp + 3 * S, gS);
          }
          leapfrog(tg, e, nunroll, m, q, p, g, t);
          if (PARK) {
            unpark(cs, leafp + 1 * S); unpark(qS, leafp + 2 * S);
            unpark(gS, leafp + 3 * S);
          }
          velocity(v, m, p);
          energy = t - half_kinetic(m, p);
        }
        ++lf_sub;
        if ((i & 1) == 0) {
          // store (velocity, momentum sum before this step) in slot
          // popcount(i)
          const int slot = __popc(i);
          T* mp = memP + (size_t)slot * K * 64;
          T* mc = memC + (size_t)slot * K * 64;
#pragma unroll
          for (int k = 0; k < K; ++k) {
            mp[k * 64 + lane] = v[k];
            mc[k * 64 + lane] = cs[k];
          }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) cs[k] = cs[k] + p[k];
        bool no_u = true;
        if (i & 1) {
          // generalized U-turn against slots [popc(i) - tz(i+1), popc(i))
          // (has_not_u_turn_at_all_index, nuts.py:949-1010)
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
          // Two checks per round, their reductions interleaved; the result
          // is the AND of the same per-slot tests (a lone last check is
          // evaluated twice).
          const int top = __popc(i);
          const int lo = top - __builtin_ctz((unsigned)(i + 1));
          for (int j = lo; j < top && no_u; j += 2) {
            const int j1 = j + 1 < top ? j + 1 : j;
            const T* mp0 = memP + (size_t)j * K * 64;
            const T* mc0 = memC + (size_t)j * K * 64;
            const T* mp1 = memP + (size_t)j1 * K * 64;
            const T* mc1 = memC + (size_t)j1 * K * 64;
            T sl0 = T(0), sr0 = T(0), sl1 = T(0), sr1 = T(0);
#pragma unroll
            for (int k = 0; k < K; ++k) {
              const T d0 = cs[k] - mc0[k * 64 + lane];
              const T d1 = cs[k] - mc1[k * 64 + lane];
              sl0 += d0 * mp0[k * 64 + lane];
              sr0 += d0 * v[k];
              sl1 += d1 * mp1[k * 64 + lane];
              sr1 += d1 * v[k];
            }
            T dl0, dr0, dl1, dr1;
            wave_sum4(sl0, sr0, sl1, sr1, dl0, dr0, dl1, dr1);
            no_u = (dl0 >= T(0)) && (dr0 >= T(0)) && (dl1 >= T(0)) &&
                   (dr1 >= T(0));
          }
        }
        if (__builtin_isnan(energy)) energy = neg_inf();
        const double de = energy - h0;
        const bool not_div = -de < max_de;
        const double wsum = log_add_exp<T>(wS, de);
        const double thresh = de - wsum;
        if (readlane_d(lu_buf, li) <= thresh) {
          copy(qS, q);
          copy(gS, g);
          tS = t;
          eS = energy;
        }
        wS = wsum;
        // energy_diff_sum counts the steps that continue
        de_buf = lane == li ? de : de_buf;
        de_ok = lane == li ? not_div : de_ok;
        nd = nd && not_div;
        csub = not_div && no_u;
        if (li == 63 || !csub || i + 1 == nsteps) {
          eds_sub += wave_sum(de_ok ? exp(fmin(de_buf, 0.0)) : 0.0);
          de_ok = false;
        }
        if (!csub) break;
      }
      if (PARK) {
        unpark(pL, tree + 0 * S); unpark(qL, tree + 1 * S);
        unpark(gL, tree + 2 * S); unpark(pR, tree + 3 * S);
        unpark(qR, tree + 4 * S); unpark(gR, tree + 5 * S);
        unpark(qC, tree + 6 * S); unpark(gC, tree + 7 * S);
        unpark(q0, tree + 8 * S); unpark(g0, tree + 9 * S);
        unpark(rho, tree + 10 * S);
      }
      eds += eds_sub;
      lf += lf_sub;
      const double tree_w = csub ? wS : neg_inf();
      const double wsum = log_add_exp<T>(tree_w, wC);
      double thresh = tree_w - wC;
      if (__builtin_isnan(thresh)) thresh = 0.0;
      const bool choose = ((double)log1m(tree_u) <= thresh) && csub;
      if (choose) {
        copy(qC, qS);
        copy(gC, gS);
        tC = tS;
        eC = eS;
      }
      wC = wsum;
      if (dir) {
        copy(qR, q); copy(pR, p); copy(gR, g); tR = t;
      } else {
        copy(qL, q); copy(pL, p); copy(gL, g); tL = t;
      }
#pragma unroll
      for (int k = 0; k < K; ++k) rho[k] = rho[k] + cs[k];
      T vL[K], vR[K];
      velocity(vL, m, pL);
      velocity(vR, m, pR);
      const T dl = dot(rho, vL);
      const T dr = dot(rho, vR);
      is_acc = is_acc || choose;
      cont = csub && (dl >= T(0)) && (dr >= T(0));
      notdiv = nd;
    }
    copy(q0, qC);
    copy(g0, gC);
    // the log-prob crosses iterations in the state dtype, exactly as it does
    // between separate one_step launches
    t0 = (double)(T)tC;
    if (lane == 0 && wic == 0) {
      const size_t o = (size_t)(a.out_row0 + it) * a.n_chains + c;
      put<T>(a.out_log_accept_ratio, o, (T)log(eds / (double)lf));
      if (a.out_leapfrogs) a.out_leapfrogs[o] = lf * nunroll;
      if (a.out_is_accepted) a.out_is_accepted[o] = is_acc;
      if (a.out_reach_max_depth) a.out_reach_max_depth[o] = cont;
      if (a.out_has_divergence) a.out_has_divergence[o] = !notdiv;
      put<T>(a.out_energy, o, (T)(eC + kconst));
      if (a.out_tree_depth) a.out_tree_depth[o] = depth;
      put<T>(a.out_target_log_prob, o, (T)t0);
    }
    if (wic == 0) trace_store(stream, a, id, dim, it, q0);
  }
  if (wic != 0) return;
  store_row((T*)a.state + row, q0, dim);
  store_row((T*)a.grad + row, g0, dim);
  if (lane == 0) ((T*)a.tlp)[crow] = (T)t0;
}

// GLM chains are bound by loads in flight, so their kernel is built for two
// waves per SIMD (the register budget that allows) and blocks of up to
// eight waves (one chain shared by a whole CU); the other families keep one
// wave per SIMD and every register for the leaf's latency.
template <typename T, int K, int FAM>
__global__ void __launch_bounds__(FAM == kFamilyGlm ? 512 : 256)
    __attribute__((amdgpu_waves_per_eu(FAM == kFamilyGlm ? 2 : 1)))
    nuts_kernel(hipmc_run_args a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  const int max_depth = a.max_tree_depth;
  // GLM chains may be shared by the W waves of a block (each evaluates a
  // share of the rows, the partial sums meet in LDS, every wave runs the
  // same control flow on the same values); otherwise a wave per chain.
  const int W = waves_per_chain<FAM>(a);
  const int c = W > 1 ? xcd_block(blockIdx.x, gridDim.x)
                      : xcd_block(blockIdx.x, gridDim.x) * waves + wave;
  const int wic = W > 1 ? wave : 0;
  if (c >= a.n_chains) return;  // whole wave (block, if shared) leaves
  const size_t cl = nuts_chain_lds_elems<T, K, FAM>(max_depth);
  const size_t wl = nuts_wave_lds_elems<T, K, FAM>();
  const size_t scr = target_scratch_elems<T, K, FAM>();
  // W > 1: [chain block][W wave blocks][partial-sum halves];
  // W == 1: one [chain block, wave block] per wave
  T* chain_base = W > 1 ? (T*)smem_raw : (T*)smem_raw + (size_t)wave * (cl + wl);
  T* wave_base = W > 1 ? (T*)smem_raw + cl + (size_t)wic * wl : chain_base + cl;
  T* memP = chain_base;  // one copy per chain; the waves of a chain write equal values
  T* memC = chain_base + (size_t)max_depth * K * 64;
  T* tree = chain_base + (size_t)2 * max_depth * K * 64;
  T* scratch = wave_base;
  T* leafp = wave_base + scr;

  const ChainId id = chain_of(a, c);
  const hipmc_problem& pb = a.problems[id.prob];
  Target<T, K, FAM> tg;
  tg.load(pb, scratch);
  if (W > 1) tg.set_split(wic, W, (unsigned char*)((T*)smem_raw + cl + (size_t)W * wl));
  const bool u1 = a.unrolled_leapfrog_steps == 1;
  if (FAM == kFamilyHier && u1 && tg.sp0) {
    nuts_chain<T, K, FAM, kBijSlot0Softplus, true>(a, tg, id, c, wic, W, memP,
                                                    memC, tree, leafp, pb);
  } else if (u1) {
    nuts_chain<T, K, FAM, kBijAny, true>(a, tg, id, c, wic, W, memP, memC,
                                         tree, leafp, pb);
  } else {
    nuts_chain<T, K, FAM, kBijAny, false>(a, tg, id, c, wic, W, memP, memC,
                                          tree, leafp, pb);
  }
}

// Chain groups (GLM): G chains of one problem per block, chain j's NUTS run
// by wave j ("owner"), every gradient a round in which all W waves of the
// block pass over X once for all G chains (Target<..., G>::group_round).
// Rounds hold two block barriers; an owner whose chain has finished, and
// every other wave, keep joining rounds (group_finish) until all owners are
// done, so every wave reaches the same barriers.  The chains of a block must
// belong to one problem (the host lays the launch out so).
// LDS: per chain its U-turn memory and parked tree vectors, per owner its
// parked leaf vectors, then the group's shared buffers (GroupLayout).
template <typename T, int K>
__host__ __device__ constexpr size_t nuts_group_lds_bytes(int max_depth, int g,
                                                         int w) {
  return ((size_t)g * (nuts_chain_lds_elems<T, K, kFamilyGlm>(max_depth) +
                       (size_t)kLeafSlots * K * 64)) * sizeof(T) +
         GroupLayout<T, K>::bytes(g, w);
}

template <typename T, int K, int G>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2)))
    nuts_group_kernel(hipmc_run_args a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wave = threadIdx.x >> 6;
  const int W = blockDim.x >> 6;
  const int max_depth = a.max_tree_depth;
  const int c0 = xcd_block(blockIdx.x, gridDim.x) * G;
  const size_t cl = nuts_chain_lds_elems<T, K, kFamilyGlm>(max_depth);
  const size_t ll = (size_t)kLeafSlots * K * 64;
  T* base = (T*)smem_raw;
  unsigned char* shared = smem_raw + (size_t)G * (cl + ll) * sizeof(T);
  const int j = wave < G ? wave : -1;
  const bool active = j >= 0 && c0 + j < a.n_chains;
  // the group's problem: that of its first chain (every chain of the block
  // shares it)
  const ChainId id0 = chain_of(a, c0);
  const hipmc_problem& pb = a.problems[id0.prob];
  Target<T, K, kFamilyGlm, G> tg;
  tg.load(pb, nullptr);
  tg.set_group(j, wave, W, shared);
  if (j >= 0 && lane_id() == 0)
    ((int*)(shared + GroupLayout<T, K>::done(G)))[j] = active ? 0 : 1;
  if (active) {
    T* cb = base + (size_t)j * (cl + ll);
    T* memP = cb;
    T* memC = cb + (size_t)max_depth * K * 64;
    T* tree = cb + (size_t)2 * max_depth * K * 64;
    T* leafp = cb + cl;
    const int c = c0 + j;
    const ChainId id = chain_of(a, c);
    nuts_chain<T, K, kFamilyGlm, kBijAny, true, G>(a, tg, id, c, 0, 1, memP,
                                                   memC, tree, leafp, pb);
  }
  tg.group_finish();
}

// ------------------------------------------------------------------ HMC
template <typename T, int K, int FAM>
__host__ __device__ constexpr size_t hmc_lds_elems() {
  return target_scratch_elems<T, K, FAM>();
}

// CALIBRATED=true: full MH(HMC) step; false: proposal only (n_iters == 1).
template <typename T, int K, int FAM, bool CALIBRATED>
__global__ void __launch_bounds__(256) hmc_kernel(hipmc_run_args a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int lane = lane_id();
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  // GLM chains may be shared by the W waves of a block (as in nuts_kernel)
  const int W = waves_per_chain<FAM>(a);
  const int c = W > 1 ? xcd_block(blockIdx.x, gridDim.x)
                      : xcd_block(blockIdx.x, gridDim.x) * waves + wave;
  const int wic = W > 1 ? wave : 0;
  const bool writer = wic == 0;  // the wave that stores a shared chain's outputs
  if (c >= a.n_chains) return;
  const size_t scr = hmc_lds_elems<T, K, FAM>();
  T* scratch = (T*)smem_raw + (size_t)wave * scr;
  const ChainId id = chain_of(a, c);
  const hipmc_problem& pb = a.problems[id.prob];
  Target<T, K, FAM> tg;
  tg.load(pb, scratch);
  if (W > 1) tg.set_split(wic, W, (unsigned char*)((T*)smem_raw + (size_t)W * scr));
  const int dim = pb.dim;
  const int crow = a.chain_row ? a.chain_row[c] : c;
  const size_t row = (size_t)crow * a.dim_stride;
  const int L = a.num_leapfrog_steps;

  T m[K];
  double kconst;
  const bool pre = load_precision(a, id, dim, m, kconst);
  T q0[K], g0[K];
  load_row(q0, (const T*)a.state + row, dim);
  load_row(g0, (const T*)a.grad + row, dim);
  double t0 = (double)((const T*)a.tlp)[crow];
  const T eps = load_step<T>(a, c);

  const Stream stream = stream_of(a, id.prob);
  for (int it = 0; it < stream.n_iters; ++it) {
    const Seed sd = iter_seed(stream, it);
    // MH splits the step seed into (proposal, acceptance)
    // (metropolis_hastings.py:183); the uncalibrated kernel alone uses the
    // seed it is given as its proposal seed.
    const Seed prop = CALIBRATED ? split(sd, 0) : sd;
    T p[K], q[K], g[K];
    draw_momentum(prop, id, dim, p);
    if (pre) precondition_momentum(m, p);
    T p_init[K];
    copy(p_init, p);
    const double ke0 = half_kinetic(m, p);
    copy(q, q0);
    copy(g, g0);
    double t = t0;
    leapfrog(tg, eps, L, m, q, p, g, t);
    const double ke1 = half_kinetic(m, p);
    // _compute_log_acceptance_correction: 0.5*safe_sum([sum p0^2, -sum p1^2])
    double corr = ke0 - ke1;
    if (!__builtin_isfinite(corr)) corr = neg_inf();
    const size_t o = (size_t)(a.out_row0 + it) * a.n_chains + c;
    if (CALIBRATED) {
      const Seed acc = split(sd, 1);
      double lar = (t - t0) + corr;  // safe_sum([t1, -t0, corr])
      if (!__builtin_isfinite(lar)) lar = neg_inf();
      const U4 uw = philox(U4{0u, (uint32_t)id.local, id.key,
                              kPurposeMHAccept},
                           acc.k0, acc.k1);
      const double u = (double)uniform_from<T>(uw.x, uw.y);
      const bool accept = log(u) < lar;
      if (it == stream.n_iters - 1 && writer) {
        if (a.out_proposed_state) store_row((T*)a.out_proposed_state + row, q, dim);
        if (a.out_initial_momentum) store_row((T*)a.out_initial_momentum + row, p_init, dim);
        if (a.out_final_momentum) store_row((T*)a.out_final_momentum + row, p, dim);
        if (lane == 0) put<T>(a.out_proposed_tlp, crow, (T)t);
      }
      if (accept) {
        copy(q0, q);
        copy(g0, g);
        t0 = (double)(T)t;
      }
      if (lane == 0 && writer) {
        put<T>(a.out_log_accept_ratio, o, (T)lar);
        if (a.out_is_accepted) a.out_is_accepted[o] = accept;
        put<T>(a.out_energy, o, (T)corr);
        put<T>(a.out_target_log_prob, o, (T)t0);
      }
    } else {
      copy(q0, q);
      copy(g0, g);
      t0 = (double)(T)t;
      if (writer && a.out_initial_momentum) store_row((T*)a.out_initial_momentum + row, p_init, dim);
      if (writer && a.out_final_momentum) store_row((T*)a.out_final_momentum + row, p, dim);
      if (lane == 0 && writer) {
        put<T>(a.out_energy, o, (T)corr);
        put<T>(a.out_target_log_prob, o, (T)t0);
      }
    }
    if (writer) trace_store(stream, a, id, dim, it, q0);
  }
  if (!writer) return;
  store_row((T*)a.state + row, q0, dim);
  store_row((T*)a.grad + row, g0, dim);
  if (lane == 0) ((T*)a.tlp)[crow] = (T)t0;
}

// ------------------------------------------------------------------ eval
template <typename T, int K, int FAM>
__global__ void __launch_bounds__(256)
    eval_kernel(const hipmc_problem* problems, const int32_t* chain_problem,
                int n_chains, int dim_stride, const T* x, T* out_tlp,
                T* out_grad) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  const int c = blockIdx.x * waves + wave;
  if (c >= n_chains) return;
  T* scratch = (T*)smem_raw + (size_t)wave * hmc_lds_elems<T, K, FAM>();
  const int prob = chain_problem ? chain_problem[c] : 0;
  const hipmc_problem& pb = problems[prob];
  Target<T, K, FAM> tg;
  tg.load(pb, scratch);
  const size_t row = (size_t)c * dim_stride;
  T q[K], g[K];
  load_row(q, x + row, pb.dim);
  const double t = tg.eval(q, g);
  if (out_grad) store_row(out_grad + row, g, pb.dim);
  if (lane_id() == 0 && out_tlp) out_tlp[c] = (T)t;
}

}  // namespace hipmc
