/*
 * emc -- the EMC reconstruction driver (emc.c:712-1088 in the reference).
 *
 * Same structure as the reference's main(): build the rotation table and the
 * detector coordinates, load the patterns and the mask, initialise the
 * model, allocate device buffers through the cuda_allocate_* entry points,
 * then iterate expand -> likelihood -> responsibilities -> scaling ->
 * M-step/insertion -> normalisation through the emc.h C-ABI, writing
 * likelihood.data (sum P log P per iteration), scale_error.data (RMS error of
 * the mean-normalised scaling factors against the true ones) and the model.
 *
 * Files, as the reference's (libspimage's HDF5 layout, written and read by
 * this library's own HDF5 code, emc_io.h):
 *   - in: with --data, patterns from printf(PATTERN, i) (the reference's
 *     Data/image_%04d.h5) and the mask from --mask (mask.h5), both sampled
 *     with --read-stride about their centre and the mask cropped to a circle
 *     (emc.c:777-864); --model-file starts from a model file (emc.c:915-928);
 *   - out: model_init.h5 and model_init_weight.h5 (emc.c:929-950),
 *     model_%04d.h5 every --save-every iterations (emc.c:1072-1082: the host
 *     model the last cuda_update_slices copied back, i.e. the accumulators
 *     before normalisation, with the host weight's mask, exactly as the
 *     reference writes them), likelihood.data, scale_error.data;
 *   - with --viewer-output, what the reference's viewer reads (viewer/):
 *     state.h5 ("iteration"), best_scaling.h5 ("scaling" [iterations][N]),
 *     average_resp_%04d.h5 ("data" [R]), best_quaternion_%04d.data (one
 *     quaternion per pattern), and responsabilities_%04d.h5 ("data" [N][R])
 *     and scaling_%04d.h5 ("data" [R][N]) while N * R stays below 2^26,
 *     and rotations.h5 ("rotations" [R][5]: q0 q1 q2 q3 weight, the layout
 *     viewer/calculate_rotation_table.py:15-21 reads); --rotations-file
 *     reads a table in that layout instead of sampling one.
 *
 * Differences from the reference, all additive:
 *   - the configuration comes from the command line instead of compile-time
 *     constants (emc.c:714-736,753,766-778);
 *   - without --data, patterns are synthetic (--config selects the
 *     workload), and scale_error.data compares with the true scales;
 *   - the host copy of the raw log-likelihoods (N*R floats) is made only
 *     with --dump-respons;
 *   - model_final.h5 holds the normalised model of the last iteration.
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>
#include <time.h>

#include "emc.h"
#include "emc_hip.h"
#include "emc_host.h"
#include "emc_io.h"

typedef struct {
  int side, n_blobs, n_patterns, n_rot, iterations, gap_h, gap_cross;
  float blob_radius, phi_sigma, photons, beamstop;
} preset;

/* The five benchmark workloads (synthetic). */
static const preset PRESETS[5] = {
    {32, 20, 1000, 4, 10, 0, 0, 6.0f, 0.2f, 2000.0f, 3.0f},
    {64, 40, 10000, 8, 20, 2, 0, 12.0f, 0.3f, 5000.0f, 4.0f},
    {128, 80, 50000, 10, 20, 0, 3, 24.0f, 0.3f, 10000.0f, 6.0f},
    {128, 80, 50000, 10, 20, 0, 3, 24.0f, 0.3f, 10000.0f, 6.0f},
    {256, 150, 20000, 8, 30, 0, 8, 48.0f, 0.5f, 300.0f, 12.0f},
};

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec * 1e3 + ts.tv_nsec / 1e6;
}

static void die_if_error(const char *where) {
  char buf[512];
  if (emc_hip_last_error(buf, sizeof buf)) {
    fprintf(stderr, "emc: %s: %s\n", where, buf);
    exit(1);
  }
}

static void usage(void) {
  fprintf(stderr,
          "usage: emc [--config K] [--patterns N] [--iterations I] [--sigma S]\n"
          "           [--precision fp32|fp16] [--insert atomic|deterministic]\n"
          "           [--interp nearest|trilinear] [--sampling n] [--full-table]\n"
          "           [--model random|truth] [--seed S] [--out DIR] [--dump-respons]\n"
          "           [--save-every K] [--quiet]\n"
          "           [--proposal T] [--topk K] [--coarse-n n] [--epochs E] [--lr LR]\n"
          "           [--likelihood gaussian|poisson] [--sparse]\n"
          "           [--data PATTERN] [--mask FILE] [--read-stride K] [--side S]\n"
          "           [--model random|truth|orientations] [--model-file FILE] [--viewer-output]\n"
          "           [--rotations-file FILE]\n"
          "  --data PATTERN: read pattern i from printf(PATTERN, i), e.g.\n"
          "                Data/image_%%04d.h5 (libspimage HDF5 layout), sampled on a\n"
          "                side x side grid with stride K (default 4) about its centre\n"
          "  --save-every K: model_%%04d.h5 every K iterations (default 1; 0: none)\n"
          "  --rotations-file FILE: the rotation table from FILE's \"rotations\" dataset\n"
          "                [R][5] (q0 q1 q2 q3 weight, the viewer's rotations file)\n"
          "                instead of the sampled one; --viewer-output writes\n"
          "                rotations.h5 in that layout\n"
          "  --proposal T: train the rotation-proposal network on iterations 1..T,\n"
          "                then score only its top-K coarse cells (default: 5 for\n"
          "                --config 3 and 4, else 0 = off)\n"
          "  --likelihood: poisson (count x log-slice) by default for --config 2..4,\n"
          "                gaussian (the reference's absolute form) for 0 and 1\n"
          "  --sparse: patterns held as CSR (default of --config 4)\n");
}

static void write_raw(const char *dir, const char *name, const float *v, size_t n) {
  char path[4096];
  if (snprintf(path, sizeof path, "%s/%s", dir, name) >= (int)sizeof path) return;
  FILE *f = fopen(path, "wb");
  if (!f) {
    fprintf(stderr, "emc: cannot write %s: %s\n", path, strerror(errno));
    return;
  }
  fwrite(v, sizeof(float), n, f);
  fclose(f);
}

/* A volume in libspimage's HDF5 layout with the mask weight > 0
 * (emc.c:929-940, 1072-1080). */
static void write_volume(const char *dir, const char *name, const float *v, const float *weight, int side) {
  char path[4096];
  if (snprintf(path, sizeof path, "%s/%s", dir, name) >= (int)sizeof path) return;
  const size_t n = (size_t)side * side * side;
  int *mask = malloc(sizeof(int) * n);
  for (size_t k = 0; k < n; k++) mask[k] = weight[k] > 0.0f ? 1 : 0;
  if (emc_write_spimage(path, v, mask, side, side, side)) fprintf(stderr, "emc: %s\n", emc_io_error());
  free(mask);
}

static void write_h5(const char *dir, const char *name, const char *dset, int type, int rank, long long d0,
                     long long d1, const void *data) {
  char path[4096];
  if (snprintf(path, sizeof path, "%s/%s", dir, name) >= (int)sizeof path) return;
  emc_h5_dataset s = {dset, type, rank, {d0, d1, 0}, data};
  if (emc_h5_write(path, &s, 1)) fprintf(stderr, "emc: %s\n", emc_io_error());
}

int main(int argc, char **argv) {
  int cfg = 0, N_images = -1, iterations = -1, n_rot = -1, full_table = 0, dump = 0, quiet = 0;
  int save_every = 1, model_truth = 0, model_orient = 0, prop_iters = -1, topk = 32, coarse_n = 4, epochs = 2;
  int read_stride = 4, side_arg = 0, viewer = 0; /* emc.c:777 */
  const char *data_pattern = NULL, *mask_file = NULL, *model_file = NULL, *rot_file = NULL;
  float lr = 1e-3f;
  float sigma = 0.4f; /* emc.c:734-735 */
  uint64_t seed = 0;
  const char *out = "output", *precision = NULL, *insert = "atomic", *interp = "nearest", *like_kind = NULL;
  int sparse = -1;
  for (int a = 1; a < argc; a++) {
    const char *k = argv[a];
    const char *v = a + 1 < argc ? argv[a + 1] : NULL;
#define TAKE(var, conv) do { if (!v) { usage(); return 2; } var = conv(v); a++; } while (0)
    if (!strcmp(k, "--config")) TAKE(cfg, atoi);
    else if (!strcmp(k, "--patterns")) TAKE(N_images, atoi);
    else if (!strcmp(k, "--iterations")) TAKE(iterations, atoi);
    else if (!strcmp(k, "--sampling")) TAKE(n_rot, atoi);
    else if (!strcmp(k, "--sigma")) TAKE(sigma, (float)atof);
    else if (!strcmp(k, "--seed")) TAKE(seed, (uint64_t)atoll);
    else if (!strcmp(k, "--save-every")) TAKE(save_every, atoi);
    else if (!strcmp(k, "--proposal")) TAKE(prop_iters, atoi);
    else if (!strcmp(k, "--topk")) TAKE(topk, atoi);
    else if (!strcmp(k, "--coarse-n")) TAKE(coarse_n, atoi);
    else if (!strcmp(k, "--epochs")) TAKE(epochs, atoi);
    else if (!strcmp(k, "--lr")) TAKE(lr, (float)atof);
    else if (!strcmp(k, "--precision")) { if (!v) { usage(); return 2; } precision = v; a++; }
    else if (!strcmp(k, "--insert")) { if (!v) { usage(); return 2; } insert = v; a++; }
    else if (!strcmp(k, "--interp")) { if (!v) { usage(); return 2; } interp = v; a++; }
    else if (!strcmp(k, "--out")) { if (!v) { usage(); return 2; } out = v; a++; }
    else if (!strcmp(k, "--rotations-file")) { if (!v) { usage(); return 2; } rot_file = v; a++; }
    else if (!strcmp(k, "--likelihood")) { if (!v) { usage(); return 2; } like_kind = v; a++; }
    else if (!strcmp(k, "--sparse")) sparse = 1;
    else if (!strcmp(k, "--model")) {
      if (!v) { usage(); return 2; }
      model_truth = !strcmp(v, "truth");
      model_orient = !strcmp(v, "orientations");
      a++;
    }
    else if (!strcmp(k, "--data")) { if (!v) { usage(); return 2; } data_pattern = v; a++; }
    else if (!strcmp(k, "--mask")) { if (!v) { usage(); return 2; } mask_file = v; a++; }
    else if (!strcmp(k, "--model-file")) { if (!v) { usage(); return 2; } model_file = v; a++; }
    else if (!strcmp(k, "--read-stride")) TAKE(read_stride, atoi);
    else if (!strcmp(k, "--side")) TAKE(side_arg, atoi);
    else if (!strcmp(k, "--viewer-output")) viewer = 1;
    else if (!strcmp(k, "--full-table")) full_table = 1;
    else if (!strcmp(k, "--dump-respons")) dump = 1;
    else if (!strcmp(k, "--quiet")) quiet = 1;
    else { usage(); return 2; }
#undef TAKE
  }
  if (cfg < 0 || cfg > 4) { usage(); return 2; }
  const preset P = PRESETS[cfg];
  if (N_images <= 0) N_images = P.n_patterns;
  if (iterations < 0) iterations = P.iterations;
  if (n_rot <= 0) n_rot = P.n_rot;
  if (!precision) precision = cfg >= 2 ? "fp16" : "fp32"; /* the precision form from 128^3 on */
  if (prop_iters < 0) prop_iters = cfg >= 3 ? 5 : 0;
  if (!like_kind) like_kind = cfg >= 2 ? "poisson" : "gaussian";
  if (sparse < 0) sparse = cfg == 4;
  if (prop_iters > 0 && strcmp(precision, "fp16") != 0) {
    fprintf(stderr, "emc: --proposal runs with --precision fp16\n");
    return 2;
  }
  if (emc_hip_set_option("precision", precision) || emc_hip_set_option("insert", insert) ||
      emc_hip_set_option("interp", interp) || emc_hip_set_option("likelihood", like_kind) ||
      emc_hip_set_option("sparse", sparse ? "on" : "off")) {
    fprintf(stderr, "emc: bad option value\n");
    return 2;
  }
  if (!quiet) emc_hip_set_option("verbose", "0");
  mkdir(out, 0755);

  /* rotations (emc.c:753-757), or a rotations file: dataset "rotations"
   * [R][5], q0 q1 q2 q3 weight per row (viewer/calculate_rotation_table.py:
   * 15-21 reads that layout) */
  float *qbuf = NULL;
  real *weights = NULL;
  int N_slices = 0;
  if (rot_file) {
    int rank = 0;
    long long dims[3] = {0, 0, 0};
    /* the shape first (a zero-element buffer: the call fails after
     * reporting rank and dims), then the data */
    (void)emc_h5_read(rot_file, "rotations", EMC_H5_F32, NULL, 0, &rank, dims);
    if (rank != 2 || dims[1] != 5 || dims[0] <= 0 || dims[0] > (1LL << 26)) {
      fprintf(stderr, "emc: %s: \"rotations\" must be [R][5] (%s)\n", rot_file, emc_io_error());
      return 1;
    }
    N_slices = (int)dims[0];
    float *rows = malloc(sizeof(float) * 5 * (size_t)N_slices);
    if (emc_h5_read(rot_file, "rotations", EMC_H5_F32, rows, 5LL * N_slices, NULL, NULL) != 5LL * N_slices) {
      fprintf(stderr, "emc: %s: %s\n", rot_file, emc_io_error());
      return 1;
    }
    qbuf = malloc(sizeof(float) * 4 * (size_t)N_slices);
    weights = malloc(sizeof(real) * (size_t)N_slices);
    for (int r = 0; r < N_slices; r++) {
      memcpy(&qbuf[4 * (size_t)r], &rows[5 * (size_t)r], sizeof(float) * 4);
      weights[r] = rows[5 * (size_t)r + 4];
    }
    free(rows);
  } else {
    const int nfull = emc_n_to_samples(n_rot);
    qbuf = malloc(sizeof(float) * 4 * (size_t)nfull);
    weights = malloc(sizeof(real) * (size_t)nfull);
    N_slices = emc_rotation_table(n_rot, full_table ? 0 : 1, qbuf, weights);
  }
  if (N_slices <= 0) { fprintf(stderr, "emc: rotation table failed\n"); return 1; }
  Quaternion **rotations = malloc(sizeof(Quaternion *) * (size_t)N_slices);
  for (int r = 0; r < N_slices; r++) {
    rotations[r] = malloc(sizeof(Quaternion));
    memcpy(rotations[r]->q, &qbuf[4 * (size_t)r], sizeof(float) * 4);
  }
  if (rot_file)
    printf("%d rotations from %s\n", N_slices, rot_file);
  else
    printf("%d rotations sampled (n = %d%s)\n", N_slices, n_rot, full_table ? "" : ", antipodal half");
  if (viewer) {
    /* the table as the viewer's rotations file: [R][5], q0..q3 and weight */
    float *rows = malloc(sizeof(float) * 5 * (size_t)N_slices);
    for (int r = 0; r < N_slices; r++) {
      memcpy(&rows[5 * (size_t)r], &qbuf[4 * (size_t)r], sizeof(float) * 4);
      rows[5 * (size_t)r + 4] = weights[r];
    }
    write_h5(out, "rotations.h5", "rotations", EMC_H5_F32, 2, N_slices, 5, rows);
    free(rows);
  }

  if (data_pattern && model_truth) {
    fprintf(stderr, "emc: --model truth needs synthetic data\n");
    return 2;
  }
  /* setup (emc.c:766-771), coordinates (emc.c:872-875) */
  Setup setup = {side_arg > 0 ? side_arg : P.side, 5.6e-9f, 16e-6f, 512, 0.15f};
  const int side = setup.side;
  const int N_2d = side * side;
  const size_t N_model = (size_t)side * side * side;
  sp_matrix xc = {side, side, malloc(sizeof(real) * N_2d)}, yc = {side, side, malloc(sizeof(real) * N_2d)},
            zc = {side, side, malloc(sizeof(real) * N_2d)};
  calculate_coordinates(setup, &xc, &yc, &zc);

  /* mask and patterns */
  sp_imatrix mask = {side, side, malloc(sizeof(int) * N_2d)};
  float *truth = NULL;
  float *image_data = malloc(sizeof(float) * (size_t)N_images * N_2d);
  float *intensities = malloc(sizeof(float) * N_images);
  double t0 = now_ms();
  int n_m = 0;
  if (data_pattern) {
    /* emc.c:777-864 (rescale_intensity 0: every intensity 1) */
    if (emc_read_patterns(data_pattern, N_images, side, read_stride, image_data) ||
        emc_read_mask(mask_file, side, read_stride, mask.data)) {
      fprintf(stderr, "emc: %s\n", emc_io_error());
      return 1;
    }
    for (int i = 0; i < N_images; i++) intensities[i] = 1.0f;
    for (int p = 0; p < N_2d; p++) n_m += mask.data[p] != 0;
    printf("%d patterns of %dx%d read (%s, stride %d), %d masked pixels (%.1f s)\n", N_images, side, side,
           data_pattern, read_stride, n_m, (now_ms() - t0) / 1e3);
  } else {
    n_m = emc_make_mask(side, P.beamstop, P.gap_h, P.gap_cross, 0, mask.data);
    emc_synth_params sp = {side, P.n_blobs, P.blob_radius, 1.0f, 2.0f, N_images, P.phi_sigma, P.photons, seed};
    truth = malloc(sizeof(float) * N_model);
    emc_synth_intensity(&sp, truth);
    emc_synth_patterns(&sp, truth, xc.data, yc.data, zc.data, mask.data, image_data, NULL, intensities);
    printf("%d patterns of %dx%d, %d masked pixels (synthetic data in %.1f s)\n", N_images, side, side, n_m,
           (now_ms() - t0) / 1e3);
  }
  sp_matrix *image_mats = malloc(sizeof(sp_matrix) * (size_t)N_images);
  sp_matrix **images = malloc(sizeof(sp_matrix *) * (size_t)N_images);
  for (int i = 0; i < N_images; i++) {
    image_mats[i].rows = side;
    image_mats[i].cols = side;
    image_mats[i].data = image_data + (size_t)i * N_2d;
    images[i] = &image_mats[i];
  }

  /* model init (emc.c:888-928): uniform random (model_input 0), every
   * pattern inserted at a random orientation (1), a model file (2), or the
   * true intensity of synthetic data */
  sp_3matrix model = {side, side, side, malloc(sizeof(real) * N_model)};
  sp_3matrix weight = {side, side, side, calloc(N_model, sizeof(real))};
  if (model_file) {
    int mx, my, mz;
    if (emc_spimage_shape(model_file, &mx, &my, &mz) || mx != side || my != side || mz != side ||
        emc_read_spimage(model_file, model.data, NULL, side, side, side)) {
      fprintf(stderr, "emc: model file %s: %s\n", model_file,
              emc_io_error()[0] ? emc_io_error() : "wrong size (emc.c:918-923)");
      return 1;
    }
  } else if (model_orient) {
    memset(model.data, 0, sizeof(float) * N_model);
    for (int i = 0; i < N_images; i++) {
      Quaternion qr;
      emc_quaternion_from_uniforms(emc_uniform(seed, 0x0121, i, 0), emc_uniform(seed, 0x0121, i, 1),
                                   emc_uniform(seed, 0x0121, i, 2), qr.q);
      insert_slice(&model, &weight, images[i], &mask, 1.0f, &qr, &xc, &yc, &zc);
    }
    for (size_t k = 0; k < N_model; k++) model.data[k] = weight.data[k] > 0.0f ? model.data[k] / weight.data[k] : 0.0f;
  } else if (model_truth) {
    memcpy(model.data, truth, sizeof(float) * N_model);
  } else {
    emc_synth_random_model(side, seed + 12345, model.data);
  }
  write_volume(out, "model_init.h5", model.data, weight.data, side);
  write_volume(out, "model_init_weight.h5", weight.data, weight.data, side);
  real *scaling = malloc(sizeof(real) * N_images);
  for (int i = 0; i < N_images; i++) scaling[i] = 1.0f;
  real *respons = dump ? malloc(sizeof(real) * (size_t)N_slices * N_images) : NULL;

  char path[1024];
  snprintf(path, sizeof path, "%s/scale_error.data", out);
  FILE *scale = fopen(path, "w");
  snprintf(path, sizeof path, "%s/likelihood.data", out);
  FILE *likelihood = fopen(path, "w");

  /* device buffers (emc.c:961-985) */
  real *slices = NULL, *d_model = NULL, *d_weight = NULL, *d_rotations = NULL, *d_x = NULL, *d_y = NULL,
       *d_z = NULL, *d_images = NULL, *d_respons = NULL, *d_scaling = NULL;
  int *d_mask = NULL;
  cuda_allocate_slices(&slices, setup, N_slices);
  cuda_allocate_model(&d_model, &model);
  cuda_allocate_model(&d_weight, &weight);
  cuda_allocate_mask(&d_mask, &mask);
  cuda_allocate_rotations(&d_rotations, rotations, N_slices);
  cuda_allocate_coords(&d_x, &d_y, &d_z, &xc, &yc, &zc);
  cuda_allocate_images(&d_images, images, N_images);
  /* N_slices*N_images*sizeof(real) bytes: the reference's int argument
   * (emc.c:983) cannot hold it beyond 2 GiB, so use the size_t form */
  emc_hip_allocate_bytes((void **)&d_respons, sizeof(real) * (size_t)N_slices * N_images);
  cuda_allocate_scaling(&d_scaling, N_images);
  die_if_error("allocation");

  if (prop_iters > 0 && emc_hip_proposal_init(qbuf, N_slices, coarse_n, P.side, topk, 256, seed) != 0)
    die_if_error("proposal setup");
  double t_total = 0;
  int *best = viewer ? malloc(sizeof(int) * N_images) : NULL;
  float *avg = viewer ? malloc(sizeof(float) * N_slices) : NULL;
  float *scale_hist = viewer ? malloc(sizeof(float) * (size_t)N_images * (iterations > 0 ? iterations : 1)) : NULL;
  const int viewer_full = viewer && (size_t)N_images * N_slices <= (1u << 26);
  float *pbuf = viewer_full ? malloc(sizeof(float) * (size_t)N_images * N_slices) : NULL;
  float *tbuf = viewer_full ? malloc(sizeof(float) * (size_t)N_images * N_slices) : NULL;
  for (int iteration = 0; iteration < iterations; iteration++) {
    const double t_it = now_ms();
    real mx = cuda_model_max(d_model, (int)N_model);
    if (!quiet) printf("model max = %g\niteration %d\n", mx, iteration);
    cuda_get_slices(&model, d_model, slices, d_rotations, d_x, d_y, d_z, N_slices);
    double t_i = now_ms();
    cuda_calculate_responsabilities(slices, d_images, d_mask, sigma, d_scaling, d_respons, N_2d, N_images,
                                    N_slices, respons);
    cuda_normalize_responsabilities(d_respons, N_slices, N_images);
    double t_e = now_ms();
    if (!quiet) printf("Expansion time = %fms\n", t_e - t_i);
    t_i = now_ms();
    const size_t nr = (size_t)N_images * N_slices;
    real total_respons = nr < 2147483648ull ? cuda_total_respons(d_respons, respons, (int)nr)
                                            : emc_hip_total_respons_n(d_respons, nr);
    fprintf(likelihood, "%g\n", total_respons);
    fflush(likelihood);
    if (!quiet) printf("likelihood = %g\n", total_respons);
    if (viewer) {
      /* viewer/rotations_module.py:43-47,116-117, viewer/slice_module.py:42 */
      if (emc_hip_respons_summary(d_respons, N_slices, N_images, best, avg)) die_if_error("respons summary");
      snprintf(path, sizeof path, "average_resp_%04d.h5", iteration);
      write_h5(out, path, "data", EMC_H5_F32, 1, N_slices, 0, avg);
      snprintf(path, sizeof path, "%s/best_quaternion_%04d.data", out, iteration);
      FILE *fq = fopen(path, "w");
      if (fq) {
        for (int i = 0; i < N_images; i++) {
          const float *qq = &qbuf[4 * (size_t)best[i]];
          fprintf(fq, "%g %g %g %g\n", qq[0], qq[1], qq[2], qq[3]);
        }
        fclose(fq);
      }
      if (viewer_full) {
        emc_hip_memcpy_to_host(pbuf, d_respons, sizeof(float) * (size_t)N_images * N_slices);
        for (int r = 0; r < N_slices; r++)
          for (int i = 0; i < N_images; i++) tbuf[(size_t)i * N_slices + r] = pbuf[(size_t)r * N_images + i];
        snprintf(path, sizeof path, "responsabilities_%04d.h5", iteration);
        write_h5(out, path, "data", EMC_H5_F32, 2, N_images, N_slices, tbuf);
      }
    }
    cuda_reset_model(&model, d_model);
    cuda_reset_model(&weight, d_weight);
    cuda_update_scaling(d_images, slices, d_mask, d_respons, d_scaling, N_images, N_slices, N_2d, scaling);
    double scale_sum = 0.0;
    for (int i = 0; i < N_images; i++) scale_sum += scaling[i];
    scale_sum /= N_images;
    double err = 0.0;
    for (int i = 0; i < N_images; i++) err += pow(scaling[i] / scale_sum - intensities[i], 2);
    err = sqrt(err / N_images);
    if (!quiet) printf("scaling error = %g\n", err);
    fprintf(scale, "%g\n", err);
    fflush(scale);
    if (viewer) {
      /* viewer/scaling_module.py:19-20, viewer/slice_module.py:77,151 */
      memcpy(scale_hist + (size_t)iteration * N_images, scaling, sizeof(float) * N_images);
      write_h5(out, "best_scaling.h5", "scaling", EMC_H5_F32, 2, iteration + 1, N_images, scale_hist);
      if (viewer_full) {
        for (int r = 0; r < N_slices; r++) memcpy(tbuf + (size_t)r * N_images, scaling, sizeof(float) * N_images);
        snprintf(path, sizeof path, "scaling_%04d.h5", iteration);
        write_h5(out, path, "data", EMC_H5_F32, 2, N_slices, N_images, tbuf);
      }
    }
    real overall = cuda_update_slices(d_images, slices, d_mask, d_respons, d_scaling, N_images, N_slices, N_2d,
                                      &model, d_model, d_x, d_y, d_z, d_rotations, weights, d_weight, setup, images);
    t_e = now_ms();
    if (!quiet) printf("Maximize time = %fms (total respons %g)\n", t_e - t_i, overall);
    t_i = now_ms();
    cuda_normalize_model(&model, d_model, d_weight);
    t_e = now_ms();
    if (!quiet) printf("Compression time = %fms\n", t_e - t_i);
    die_if_error("iteration");
    if (iteration < prop_iters) {
      /* the proposal network learns from this iteration's responsibilities;
       * after the last training iteration it predicts every pattern's cells */
      double loss = 0.0;
      emc_hip_proposal_train(d_images, d_mask, N_2d, d_respons, N_images, N_slices, epochs, lr, &loss);
      if (!quiet) printf("proposal training loss = %g\n", loss);
      if (iteration == prop_iters - 1) {
        emc_hip_proposal_predict(d_images, d_mask, N_2d, N_images);
        emc_hip_set_option("proposal", "on");
      }
      die_if_error("proposal");
    } else if (prop_iters > 0) {
      double st[13];
      emc_hip_proposal_stats(st, 13);
      if (!quiet)
        printf("proposal: %.3g pairs scored, %.0f patterns on the whole table, %.0f of %.0f audited below "
               "min mass (min captured %.6f)\n", st[6], st[11], st[3], st[2], st[4]);
    }
    const double dt = now_ms() - t_it;
    t_total += dt;
    printf("iteration %d: %.1f ms, %.3g pattern x rotation pairs/s, likelihood %g, scaling error %g\n", iteration,
           dt, (double)N_images * N_slices / (dt / 1e3), total_respons, err);
    if (save_every > 0 && (iteration % save_every == 0 || iteration == iterations - 1)) {
      /* emc.c:1072-1082: the host model holds the accumulators before
       * normalisation (emc_cuda.cu:352 copies it before cuda_normalize_model),
       * the mask is the host weight's, which the iterations do not update */
      snprintf(path, sizeof path, "model_%04d.h5", iteration);
      write_volume(out, path, model.data, weight.data, side);
    }
    if (viewer) {
      /* viewer/viewer.py:14-21,96-97 */
      const int it = iteration;
      write_h5(out, "state.h5", "iteration", EMC_H5_I32, 0, 0, 0, &it);
    }
    if (respons) {
      snprintf(path, sizeof path, "respons_%04d.raw", iteration);
      write_raw(out, path, respons, (size_t)N_slices * N_images);
    }
  }
  /* final normalised model (and the device weight as its mask) */
  float *final_model = malloc(sizeof(float) * N_model);
  float *final_weight = malloc(sizeof(float) * N_model);
  emc_hip_memcpy_to_host(final_model, d_model, sizeof(float) * N_model);
  emc_hip_memcpy_to_host(final_weight, d_weight, sizeof(float) * N_model);
  write_volume(out, "model_final.h5", final_model, final_weight, side);
  if (truth) write_volume(out, "truth.h5", truth, truth, side);
  printf("%d iterations in %.1f s\n", iterations, t_total / 1e3);
  fclose(scale);
  fclose(likelihood);
  die_if_error("end");
  emc_hip_shutdown();
  return 0;
}
