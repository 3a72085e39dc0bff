// Command-line driver of the oracle (test infrastructure): same flags as the reference solver,
// writes diagnostics.csv, profiles_timeavg.csv and snapshots into --outdir.
#include "nse2d_oracle.hpp"

int main(int argc, char** argv) { return orc_run_main(argc - 1, argv + 1); }
