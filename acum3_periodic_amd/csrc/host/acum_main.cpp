// acum -- drop-in driver for the reference program (main.cu:11-286): reads `input`, obs.dat,
// surfaces/surface_k.dat (and input_BB, BB_inputs/* when BB_noise = 1) from a run directory,
// preprocesses on the host, runs the tonal (and broadband) paths on the GPU through the C-ABI and
// writes pT.out, pL.out, pA.out (and bb_oaspl.out, bb_spl.dat) into the same directory.
//
//   acum [run_dir] [--fp64 | --mixed] [--device N] [--gpu-prep] [--cache]
//   acum --generate CONFIG OUT_DIR      (writes a synthetic case: CONFIG 1, 2, 3 or 5)
// --gpu-prep preprocesses on the GPU in place (acum_run_stage: the element data go to the device
// once and never come back); --cache writes binary caches of the parsed surface files, which
// later runs read instead of the text.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include "../../../include/acum.h"

static double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

int main(int argc, char** argv) {
  std::string dir = ".";
  int precision = ACUM_MIXED, device = 0;
  bool gpu_prep = false, cache = false;
  for (int i = 1; i < argc; i++) {
    const std::string a = argv[i];
    if (a == "--fp64") precision = ACUM_FP64;
    else if (a == "--mixed") precision = ACUM_MIXED;
    else if (a == "--device" && i + 1 < argc) device = std::atoi(argv[++i]);
    else if (a == "--gpu-prep") gpu_prep = true;
    else if (a == "--cache") cache = true;
    else if (a == "--generate" && i + 2 < argc) {
      acum_synth_params p;
      if (acum_synth_defaults(std::atoi(argv[i + 1]), &p)) {
        std::fprintf(stderr, "acum: %s\n", acum_host_error());
        return 2;
      }
      acum_run* r = nullptr;
      if (acum_synth(&p, &r) || acum_run_save(r, argv[i + 2])) {
        std::fprintf(stderr, "acum: %s\n", acum_host_error());
        return 2;
      }
      acum_run_free(r);
      return 0;
    } else if (a == "-h" || a == "--help") {
      std::printf("usage: acum [run_dir] [--fp64|--mixed] [--device N] [--gpu-prep] [--cache]\n"
                  "       acum --generate CONFIG OUT_DIR\n");
      return 0;
    } else dir = a;
  }
  double t0 = now_ms();
  acum_run* run = nullptr;
  if (acum_run_load(dir.c_str(), &run)) {
    std::fprintf(stderr, "acum: %s\n", acum_host_error());
    return 1;
  }
  double t1 = now_ms();
  if (cache && acum_run_save_cache(run, dir.c_str())) {
    std::fprintf(stderr, "acum: %s\n", acum_host_error());
    return 1;
  }
  if (!gpu_prep && acum_run_preprocess(run)) {
    std::fprintf(stderr, "acum: %s\n", acum_host_error());
    return 1;
  }
  double t2 = now_ms();
  acum_inputs in;
  acum_run_inputs(run, &in);
  acum_ctx* ctx = nullptr;
  if (acum_ctx_create(&ctx, device)) {
    std::fprintf(stderr, "acum: no usable GPU (device %d)\n", device);
    return 1;
  }
  acum_tonal_case tc;
  (void)acum_run_tonal_case(run, precision, &tc);  // sizes; not preprocessed yet with --gpu-prep
  const size_t n = (size_t)tc.nTime * tc.nObs;
  std::vector<float> pT(n), pL(n), pA(n);
  int rc;
  if (gpu_prep) {
    int32_t h = 0;
    rc = acum_run_stage(ctx, run, &h);
    if (!rc) rc = acum_tonal_run(ctx, h, precision, 1, 0);
    if (!rc) rc = acum_tonal_fetch(ctx, h, 0, pT.data(), pL.data(), pA.data());
    if (!rc) rc = acum_tonal_release(ctx, h);
  } else {
    rc = acum_tonal(ctx, &tc, pT.data(), pL.data(), pA.data());
  }
  if (rc) {
    std::fprintf(stderr, "acum: tonal: %s\n", acum_last_error(ctx));
    return 1;
  }
  acum_timing tt;
  acum_get_timing(ctx, &tt);
  double t3 = now_ms();
  std::vector<float> spl, oaspl;
  if (in.BB_noise == 1) {
    acum_bpm_case bc;
    acum_run_bpm_case(run, &bc);
    spl.assign((size_t)bc.nObs * bc.nf2 * 7, -100.0f);
    oaspl.assign(bc.nObs, -100.0f);
    rc = acum_bpm(ctx, &bc, spl.data(), oaspl.data());
    if (rc) {
      std::fprintf(stderr, "acum: broadband: %s\n", acum_last_error(ctx));
      return 1;
    }
  }
  double t4 = now_ms();
  if (acum_run_write_tonal(run, dir.c_str(), pT.data(), pL.data(), pA.data()) ||
      (in.BB_noise == 1 && acum_run_write_bb(run, dir.c_str(), oaspl.data(), spl.data()))) {
    std::fprintf(stderr, "acum: %s\n", acum_host_error());
    return 1;
  }
  double t5 = now_ms();
  std::printf("read %.1f ms, preprocess %.1f ms, tonal %.1f ms (kernels %.2f ms, %.3g evals), broadband %.1f ms, write %.1f ms\n",
              t1 - t0, t2 - t1, t3 - t2, tt.eval_ms + tt.reduce_ms, (double)tt.evals, t4 - t3, t5 - t4);
  acum_ctx_destroy(ctx);
  acum_run_free(run);
  return 0;
}
