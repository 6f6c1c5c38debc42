// lbm_solve — drop-in replacement for the reference's serial executable
// (serial/main.c): reads ./SetUpData.ini, ./Mesh/D2node.dat and
// ./Mesh/BCconnectors.dat, writes ./Results/ with the same names and
// formats, exits 1 after printing DIVERGENCE! when the residual turns NaN.
//
// Without options it reproduces the reference behaviour (FP64, reference
// start, reference residual).  Options select the engine's other modes.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../include/lbm.h"

static void usage() {
  std::fprintf(stderr,
               "usage: lbm_solve [--ini FILE] [--nodes FILE] [--bc FILE] [--out DIR]\n"
               "                 [--precision f64|f32|f16|f16sr] [--init reference|rest]\n"
               "                 [--resid reference|full|velocity|none] [--resid-every K]\n"
               "                 [--trt-lambda X] [--trt-fp64] [--mrt-textbook]\n"
               "                 [--mrt-rates s0,s1,...,s8] [--restart FILE]\n"
               "                 [--checkpoint FILE] [--quiet]\n");
}

int main(int argc, char** argv) {
  const char* ini_file = "SetUpData.ini";
  const char* nodes = "Mesh/D2node.dat";
  const char* bc = "Mesh/BCconnectors.dat";
  const char* out = "";
  const char* restart = nullptr;
  const char* checkpoint = nullptr;
  int verbose = 1;
  bool custom = false;
  lbm_ini ini;
  std::string precision, init, resid;
  int resid_every = 0, trt_fp64 = 0, mrt_textbook = 0;
  double trt_lambda = 0.0;
  double rates[9];
  bool have_rates = false;
  for (int a = 1; a < argc; a++) {
    auto next = [&](void) -> const char* {
      if (a + 1 >= argc) {
        usage();
        std::exit(2);
      }
      return argv[++a];
    };
    if (!std::strcmp(argv[a], "--ini")) ini_file = next();
    else if (!std::strcmp(argv[a], "--nodes")) nodes = next();
    else if (!std::strcmp(argv[a], "--bc")) bc = next();
    else if (!std::strcmp(argv[a], "--out")) out = next();
    else if (!std::strcmp(argv[a], "--precision")) { precision = next(); custom = true; }
    else if (!std::strcmp(argv[a], "--init")) { init = next(); custom = true; }
    else if (!std::strcmp(argv[a], "--resid")) { resid = next(); custom = true; }
    else if (!std::strcmp(argv[a], "--resid-every")) { resid_every = std::atoi(next()); custom = true; }
    else if (!std::strcmp(argv[a], "--trt-lambda")) { trt_lambda = std::atof(next()); custom = true; }
    else if (!std::strcmp(argv[a], "--trt-fp64")) { trt_fp64 = 1; custom = true; }
    else if (!std::strcmp(argv[a], "--mrt-textbook")) { mrt_textbook = 1; custom = true; }
    else if (!std::strcmp(argv[a], "--mrt-rates")) {
      const char* p = next();
      for (int k = 0; k < 9; k++) {
        char* e;
        rates[k] = std::strtod(p, &e);
        if (e == p) { usage(); return 2; }
        p = (*e == ',') ? e + 1 : e;
      }
      have_rates = true;
      custom = true;
    } else if (!std::strcmp(argv[a], "--restart")) restart = next();
    else if (!std::strcmp(argv[a], "--checkpoint")) checkpoint = next();
    else if (!std::strcmp(argv[a], "--quiet")) verbose = 0;
    else { usage(); return 2; }
  }
  lbm_options opt;
  if (custom) {
    if (lbm_read_ini(ini_file, &ini) != LBM_OK) {
      std::fprintf(stderr, "%s\n", lbm_last_error());
      return 2;
    }
    lbm_options_from_ini(&ini, &opt);
    if (precision == "f32") opt.precision = LBM_PREC_F32;
    else if (precision == "f16") opt.precision = LBM_PREC_F16;
    else if (precision == "f16sr") opt.precision = LBM_PREC_F16_SR;
    else if (!precision.empty() && precision != "f64") { usage(); return 2; }
    if (init == "rest") opt.init = LBM_INIT_REST;
    else if (!init.empty() && init != "reference") { usage(); return 2; }
    if (resid == "full") opt.resid_kind = LBM_RESID_FULL;
    else if (resid == "velocity") opt.resid_kind = LBM_RESID_VELOCITY;
    else if (resid == "none") opt.resid_kind = LBM_RESID_NONE;
    else if (!resid.empty() && resid != "reference") { usage(); return 2; }
    if (resid_every > 0) opt.resid_every = resid_every;
    opt.trt_lambda = trt_lambda;
    opt.trt_fp64_parts = trt_fp64;
    opt.mrt_textbook_meq = mrt_textbook;
    if (have_rates) {
      opt.mrt_custom_rates = 1;
      for (int k = 0; k < 9; k++) opt.mrt_rates[k] = rates[k];
    }
  }
  const int rc = lbm_run_checkpointed(ini_file, nodes, bc, out, custom ? &opt : nullptr, restart,
                                      checkpoint, verbose);
  if (rc == LBM_DIVERGED) return 1;  // serial/ComputeResiduals.c:50
  return rc == LBM_OK ? 0 : 2;
}
