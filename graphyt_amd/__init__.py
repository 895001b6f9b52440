"""graphyt_amd — MI355X-native MPM time step (see include/graphyt_amd.h)."""
