// sw2d_run -- drop-in replacement of the reference executable.
//
// Like SW2D-GPU (SW2D-GPU_linux/Readme.txt:15-19) it is run from the
// directory that holds db/, reads db/dir.txt and writes each case's
// output/ files.  Optional flags: --db DIR (default "db"), --fp32,
// --cfl C (adaptive step, capped at 2*dt), --step-forcing.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "sw2d.h"

int main(int argc, char **argv) {
  std::string db = "db";
  sw2d_params p;
  sw2d_params_default(&p);
  for (int i = 1; i < argc; i++) {
    if (!std::strcmp(argv[i], "--db") && i + 1 < argc) db = argv[++i];
    else if (!std::strcmp(argv[i], "--fp32")) p.precision = SW2D_FP32;
    else if (!std::strcmp(argv[i], "--cfl") && i + 1 < argc) p.cfl = std::atof(argv[++i]);
    else if (!std::strcmp(argv[i], "--step-forcing")) p.forcing_mode = SW2D_FORCING_STEP;
    else {
      std::fprintf(stderr, "usage: %s [--db DIR] [--fp32] [--cfl C] [--step-forcing]\n", argv[0]);
      return 2;
    }
  }
  const int rc = sw2d_run_db(db.c_str(), &p);
  if (rc) {
    std::fprintf(stderr, "sw2d_run: error %d: %s\n", rc, sw2d_last_error());
    return 1;
  }
  return 0;
}
