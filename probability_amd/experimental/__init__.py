"""Experimental APIs (`tfp.experimental`): streaming reducers and running
statistics on the device."""

from probability_amd.experimental import mcmc
from probability_amd.experimental import stats
