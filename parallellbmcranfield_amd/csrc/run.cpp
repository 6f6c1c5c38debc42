// run.cpp — result writers and the whole-program driver (the reference's
// main + Iteration), writing Results/ with the reference's names and formats.
#include <sys/stat.h>
#include <sys/types.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "fmt_fixed.hpp"
#include "lbm_host.hpp"
#include "result_writer.hpp"

using namespace lbmhost;

namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::string join(const std::string& dir, const char* name) {
  if (dir.empty()) return name;
  return dir + "/" + name;
}

}  // namespace

extern "C" int lbm_solver_write_results(lbm_solver* s, const char* path, int format,
                                        int final_layout, int row0_bug) {
  if (!s || !path) return LBM_EINVAL;
  const lbm_case* c = s->cs;
  const int n = c->n, m = c->m;
  const long nc = (long)n * m;
  std::vector<double> rho((size_t)nc), u((size_t)nc), v((size_t)nc), f;
  int rc = lbm_solver_get_macro(s, rho.data(), u.data(), v.data());
  if (rc) return rc;
  if (final_layout) {
    f.resize((size_t)nc * 9);
    rc = lbm_solver_get_f(s, f.data(), nullptr);
    if (rc) return rc;
  }
  const bool csv = format != LBM_FORMAT_TECPLOT;
  char head[512];
  std::string header;
  // serial/FileIO.c:298-348 (WriteResults) and :354-427 (WriteFinalResults)
  if (!final_layout) {
    if (csv) {
      header = "x,y,u,v,vel_mag,rho,press,fluid\n";
    } else {
      header = "Title = \"LBM results\"\n";
      header += "Variables = \"x\",\"y\",\"u\",\"v\",\"u mag\",\"rho\",\"press\",\"fluid\"\n";
      std::snprintf(head, sizeof head, "Zone i=%d, j=%d, f=point\n", n, m);
      header += head;
    }
    // "%f, %f, %f, %f, %f, %f, %f, %d\n" (CSV) or space-separated (Tecplot)
    const char* sep = csv ? ", " : " ";
    return write_lines(path, header, nc, [&](long i, char* buf, size_t) {
      const double a[7] = {c->x(i), c->y(i), u[i], v[i], std::sqrt(u[i] * u[i] + v[i] * v[i]),
                           rho[i], rho[i] / 3.0};
      return fmt_row(buf, a, 7, c->fluid[i], nullptr, 0, sep);
    });
  }
  if (csv) {
    header = "x,y,u,v,vel_mag,rho,press,fluid,f0,f1,f2,f3,f4,f5,f6,f7,f8\n";
  } else {
    header = "Title = \"LBM results\"\n";
    header += "Variables = \"x\",\"y\",\"u\",\"v\",\"u mag\",\"rho\",\"press\",\"fluid\",\"f0\",\"f1\",\"f2\",\"f3\",\"f4\",\"f5\",\"f6\",\"f7\",\"f8\"\n";
    std::snprintf(head, sizeof head, "Zone i=%d, j=%d, f=point\n", n, m);
    header += head;
  }
  // "%f, %f, %f, %f, %f, %f, %f, %d, %f x 9\n" (CSV) or space-separated (Tecplot)
  const char* sep = csv ? ", " : " ";
  return write_lines(path, header, nc, [&](long r, char* buf, size_t) {
    const long q = row0_bug ? (r % n) : r;  // FileIO.c:368 indexes Cells+i for every row j
    const double a[7] = {c->x(q), c->y(q), u[q], v[q], std::sqrt(u[q] * u[q] + v[q] * v[q]),
                         rho[q], rho[q] / 3};
    return fmt_row(buf, a, 7, c->fluid[q], &f[(size_t)q * 9], 9, sep);
  });
}

// serial/Iterate.c:17-410 with serial/main.c:20-93 around it.
static int run_impl(const lbm_ini& ini, const char* node_file, const char* bc_file,
                    const char* out_dir, const lbm_options* overrides, int verbose,
                    const char* restart_file = nullptr, const char* checkpoint_file = nullptr) {
  const double t_start = now_s();
  const std::string res_dir = join(out_dir ? out_dir : "", "Results");
  if (verbose) {
    std::printf("\n###############################################\n");
    std::printf("#### This is a 2D lattice-Boltzmann solver ####\n");
    std::printf("####   MI355X (HIP) engine, one GPU        ####\n");
    std::printf("###############################################\n");
  }
  mkdir(res_dir.c_str(), S_IRWXU | S_IRGRP | S_IXGRP);  // ShellFunctions.c:164-179

  lbm_case* cs = nullptr;
  int rc = lbm_case_load(node_file ? node_file : "Mesh/D2node.dat",
                         bc_file ? bc_file : "Mesh/BCconnectors.dat", &cs);
  if (rc) {
    std::fprintf(stderr, "%s\n", lbm_last_error());
    return rc;
  }
  lbm_options opt;
  lbm_options_from_ini(&ini, &opt);
  if (overrides) opt = *overrides;

  const std::string log_path = res_dir + "/logFile.log";
  FILE* log = std::fopen(log_path.c_str(), "w");
  if (!log) {
    set_error("cannot write " + log_path);
    lbm_case_free(cs);
    return LBM_EIO;
  }
  // header (Iterate.c:113-165)
  std::fprintf(log, "This is the 2D lattice Boltzmann *.log file\n\n");
  std::fprintf(log, "\n:::: Imported variables from the *.ini file :::: \n");
  std::fprintf(log, ">>> Uavg              : %3.6f\n", ini.uavg);
  std::fprintf(log, ">>> Vavg              : %3.6f\n", ini.vavg);
  std::fprintf(log, ">>> Initial density   : %2.1f\n", ini.rho_ini);
  std::fprintf(log, ">>> Viscosity         : %3.8f\n", ini.viscosity);
  std::fprintf(log, ">>> # of iterations   : %1.1d\n", ini.iterations);
  std::fprintf(log, ">>> Autosave after    : %1.1d\n", ini.autosave_after);
  std::fprintf(log, ">>> Autosave every    : %1.1d\n", ini.autosave_every);
  std::fprintf(log, ">>> Convergence Veloc : %3.8f\n", ini.convergence_crit_veloc);
  std::fprintf(log, ">>> Convergence Rho   : %3.8f\n", ini.convergence_crit_rho);
  std::fprintf(log, ">>> Former data iter# : %1.1d\n", ini.read_former_data);
  switch (ini.collision_model) {
    case 1: std::fprintf(log, ">>> CollisionModel   : BGKW\n"); break;
    case 2: std::fprintf(log, ">>> CollisionModel   : TRT\n"); break;
    case 3: std::fprintf(log, ">>> CollisionModel   : MRT\n"); break;
  }
  switch (ini.inlet_profile) {
    case 1: std::fprintf(log, ">>> InletProfile     : ON\n"); break;
    case 2: std::fprintf(log, ">>> InletProfile     : OFF\n"); break;
  }
  switch (ini.outlet_profile) {
    case 1: std::fprintf(log, ">>> OutletProfile    : ON\n"); break;
    case 2: std::fprintf(log, ">>> OutletProfile    : OFF\n"); break;
  }
  switch (ini.curved_boundaries) {
    case 1: std::fprintf(log, ">>> CurvedBoundaries : ON\n"); break;
    case 2: std::fprintf(log, ">>> CurvedBoundaries : OFF\n"); break;
  }
  switch (ini.postproc_prog) {
    case 1: std::fprintf(log, ">>> Results format   : Paraview (*.csv)\n"); break;
    case 2: std::fprintf(log, ">>> Results format   : Tecplot (*.dat)\n"); break;
  }
  if (ini.calculate_drag_lift != 0)
    std::fprintf(log, ">>> Drag, lift @ BC  : %d\n", ini.calculate_drag_lift);
  else
    std::fprintf(log, ">>> Drag, lift was not calculated\n");
  std::fprintf(log, "\n:::: Calculated variables from mesh :::: \n");
  std::fprintf(log, ">>> Grid spacing        = %f\n", cs->delta);
  std::fprintf(log, ">>> # of nodes in x (n) = %d\n", cs->n);
  std::fprintf(log, ">>> # of nodes in y (m) = %d\n", cs->m);
  std::fprintf(log, ">>> NumInletNodes       = %d\n", cs->num_inlet);
  std::fprintf(log, ">>> MaxInletCoordY      = %f\n", cs->max_inlet_y);
  std::fprintf(log, ">>> MinInletCoordY      = %f\n", cs->min_inlet_y);

  std::fprintf(log, "\n:::: Initializing ::::\n");
  if (verbose) std::printf("\n:::: Initializing ::::\n");
  const double t_init0 = now_s();
  lbm_solver* s = nullptr;
  rc = lbm_solver_create(cs, &opt, &s);
  if (rc) {
    std::fprintf(stderr, "%s\n", lbm_last_error());
    std::fclose(log);
    lbm_case_free(cs);
    return rc;
  }
  if (restart_file) {
    rc = lbm_solver_load(s, restart_file);
    if (rc) {
      std::fprintf(stderr, "%s\n", lbm_last_error());
      std::fclose(log);
      lbm_solver_destroy(s);
      lbm_case_free(cs);
      return rc;
    }
  }
  const double t_init = now_s() - t_init0;
  std::fprintf(log, "\n:::: Initialization done! ::::\n");
  const bool csv = ini.postproc_prog != 2;
  const std::string init_out = res_dir + (csv ? "/InitialData.csv" : "/InitialData.dat");
  double t_writing = 0.0;
  if (ini.postproc_prog == 1 || ini.postproc_prog == 2) {
    rc = lbm_solver_write_results(s, init_out.c_str(), ini.postproc_prog, 0, 0);
    if (rc) goto fail;
    if (verbose) std::printf("\nInitialized data was written to %s\n", ("Results" + init_out.substr(res_dir.size())).c_str());
  }
  {
    const std::string rpath = res_dir + "/residuals.dat";
    FILE* resf = std::fopen(rpath.c_str(), "w");
    if (!resf) {
      set_error("cannot write " + rpath);
      rc = LBM_EIO;
      goto fail;
    }
    std::fprintf(resf, "Iter L2_norm L2_norm_weighted Drag Lift\n");
    std::fprintf(log, "\n:::: Start Iterations ::::\n");
    if (verbose) std::printf("\n:::: Start Iterations ::::\n");

    std::string autosave_name = "NOWHERE!";
    const double t_iter0 = now_s();
    double t_resid = 0.0, t_step = 0.0;
    std::vector<int> hi(1024);
    std::vector<double> hv(4096);
    bool diverged = false;
    for (;;) {
      const int it = lbm_solver_iter(s);
      // run to the next progress print (every 100) or autosave point
      int next = (it / 100 + 1) * 100;
      if (ini.autosave_every > 0) {
        const int a = (it / ini.autosave_every + 1) * ini.autosave_every;
        if (a < next) next = a;
      }
      const int chunk = next - it;
      int done = 0;
      const double ts = now_s();
      int src = lbm_solver_step(s, chunk, &done);
      t_step += now_s() - ts;
      const double tr = now_s();
      for (;;) {
        int cnt = 0;
        lbm_solver_residual_history(s, 1024, hi.data(), hv.data(), &cnt);
        for (int k = 0; k < cnt; k++)
          std::fprintf(resf, "%d %5.4e %5.4e %f %f\n", hi[k], hv[4 * k], hv[4 * k + 1],
                       hv[4 * k + 2], hv[4 * k + 3]);
        if (cnt < 1024) break;
      }
      t_resid += now_s() - tr;
      if (src == LBM_DIVERGED) {
        diverged = true;
        break;
      }
      if (src) {
        rc = src;
        std::fclose(resf);
        goto fail;
      }
      const int now = lbm_solver_iter(s);
      if (done > 0) {
        if (now % 100 == 0 && verbose) {
          double r[4];
          lbm_solver_residuals(s, r);
          std::printf("Iterations: %05d/%05d || ", now, ini.iterations);
          std::printf("Residuals: l2norm  %e; L2_norm_weighted  %e\n", r[0], r[1]);
        }
        // Iterate.c:316-330
        if (ini.autosave_every > 0 && now % ini.autosave_every == 0 && now > ini.autosave_after &&
            (ini.postproc_prog == 1 || ini.postproc_prog == 2)) {
          char name[64];
          std::snprintf(name, sizeof name, csv ? "autosave_iter%05d.csv" : "autosave_iter%05d.dat", now);
          const double tw = now_s();
          rc = lbm_solver_write_results(s, (res_dir + "/" + name).c_str(), ini.postproc_prog, 0, 0);
          t_writing += now_s() - tw;
          if (rc) {
            std::fclose(resf);
            goto fail;
          }
          autosave_name = std::string("Results/") + name;
        }
      }
      if (done < chunk) break;  // loop condition failed
    }
    std::fclose(resf);
    if (diverged) {
      // ComputeResiduals.c:47-51
      std::printf("\nDIVERGENCE!\n");
      std::fclose(log);
      lbm_solver_destroy(s);
      lbm_case_free(cs);
      return LBM_DIVERGED;
    }
    const double t_iteration = now_s() - t_iter0;
    const double t_overall = now_s() - t_start;
    // Iterate.c:341-371.  The step pass fuses collision, streaming, boundaries
    // and the macroscopic update, so its time is reported under Collision.
    std::fprintf(log, "\nOverall calculations took %f seconds\n", t_overall);
    std::fprintf(log, "Main while loop took %f seconds\n", t_iteration);
    std::fprintf(log, "Initialization took %f seconds\n", t_init);
    std::fprintf(log, "Collision took %f seconds\n", t_step);
    std::fprintf(log, "UpdateF took %f seconds\n", 0.0);
    std::fprintf(log, "Streaming took %f seconds\n", 0.0);
    std::fprintf(log, "Calculating Boundaries took %f seconds\n", 0.0);
    std::fprintf(log, "Update Macroscopic took %f seconds\n", 0.0);
    std::fprintf(log, "Calculating Residuals took %f seconds\n", t_resid);
    std::fprintf(log, "Writing results took %f seconds\n", t_writing);
    std::fprintf(log, "\n:::: Iterations done! ::::\n");
    std::fclose(log);
    log = nullptr;
    const std::string tpath = res_dir + "/SerialTimeMeasuerment.dat";
    FILE* tf = std::fopen(tpath.c_str(), "w");
    if (tf) {
      std::fprintf(tf, "tOverall %f\n", t_overall);
      std::fprintf(tf, "tIteration %f\n", t_iteration);
      std::fprintf(tf, "tInitialization %f\n", t_init);
      std::fprintf(tf, "tCollision %f\n", t_step);
      std::fprintf(tf, "tUpdateF %f\n", 0.0);
      std::fprintf(tf, "tStreaming %f\n", 0.0);
      std::fprintf(tf, "tBoundaries %f\n", 0.0);
      std::fprintf(tf, "tUpdateMacro %f\n", 0.0);
      std::fprintf(tf, "tResiduals %f\n", t_resid);
      std::fprintf(tf, "tWriting %f\n", t_writing);
      std::fclose(tf);
    }
    if (checkpoint_file) {
      rc = lbm_solver_save(s, checkpoint_file);
      if (rc) goto fail;
    }
    std::string final_out = res_dir + (csv ? "/FinalData.csv" : "/FinalData.dat");
    if (ini.postproc_prog == 1 || ini.postproc_prog == 2) {
      // FileIO.c:354-427, row-0 repetition included
      rc = lbm_solver_write_results(s, final_out.c_str(), ini.postproc_prog, 1, 1);
      if (rc) goto fail;
    }
    if (verbose) {
      std::printf("\n\nLog was written to %s\n", "Results/logFile.log");
      std::printf("Last autosave result can be found at %s\n", autosave_name.c_str());
      std::printf("Residuals were written to Results/residuals.dat\n");
      std::printf("Profiling results were written to Results/SerialTimeMeasuerment.dat\n");
      std::printf("Final results were written to %s\n",
                  csv ? "Results/FinalData.csv" : "Results/FinalData.dat");
    }
  }
  lbm_solver_destroy(s);
  lbm_case_free(cs);
  return LBM_OK;
fail:
  if (log) std::fclose(log);
  std::fprintf(stderr, "%s\n", lbm_last_error());
  lbm_solver_destroy(s);
  lbm_case_free(cs);
  return rc;
}

extern "C" int lbm_run(const char* ini_file, const char* node_file, const char* bc_file,
                       const char* out_dir, const lbm_options* overrides, int verbose) {
  lbm_ini ini;
  int rc = lbm_read_ini(ini_file ? ini_file : "SetUpData.ini", &ini);
  if (rc) {
    std::fprintf(stderr, "%s\n", lbm_last_error());
    return rc;
  }
  return run_impl(ini, node_file, bc_file, out_dir, overrides, verbose);
}

extern "C" int lbm_run_checkpointed(const char* ini_file, const char* node_file,
                                    const char* bc_file, const char* out_dir,
                                    const lbm_options* overrides, const char* restart_file,
                                    const char* checkpoint_file, int verbose) {
  lbm_ini ini;
  int rc = lbm_read_ini(ini_file ? ini_file : "SetUpData.ini", &ini);
  if (rc) {
    std::fprintf(stderr, "%s\n", lbm_last_error());
    return rc;
  }
  return run_impl(ini, node_file, bc_file, out_dir, overrides, verbose, restart_file,
                  checkpoint_file);
}

extern "C" int lbm_iteration(const char* NodeDataFile, const char* BCconnectorDataFile,
                             float Uavg, float Vavg, float rho_ini, float Viscosity,
                             int InletProfile, int CollisionModel, int CurvedBoundaries,
                             int OutletProfile, int Iterations, int AutosaveAfter,
                             int AutosaveEvery, int postproc_prog, int CalculateDragLift,
                             float ConvergenceCritVeloc, float ConvergenceCritRho,
                             int ReadFormerData) {
  lbm_ini ini;
  ini.uavg = Uavg;
  ini.vavg = Vavg;
  ini.rho_ini = rho_ini;
  ini.viscosity = Viscosity;
  ini.inlet_profile = InletProfile;
  ini.collision_model = CollisionModel;
  ini.curved_boundaries = CurvedBoundaries;
  ini.outlet_profile = OutletProfile;
  ini.iterations = Iterations;
  ini.autosave_every = AutosaveEvery;
  ini.autosave_after = AutosaveAfter;
  ini.postproc_prog = postproc_prog;
  ini.calculate_drag_lift = CalculateDragLift;
  ini.convergence_crit_veloc = ConvergenceCritVeloc;
  ini.convergence_crit_rho = ConvergenceCritRho;
  ini.read_former_data = ReadFormerData;
  return run_impl(ini, NodeDataFile, BCconnectorDataFile, nullptr, nullptr, 1);
}
