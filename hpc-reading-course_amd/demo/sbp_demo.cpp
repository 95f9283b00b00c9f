// sbp_demo.cpp -- PETSc-free drivers of the reference's five demos on the GPU
// path: adv_1D, adv_2D, wave, wave_hom and reflection, selected at compile
// time (-DSBP_DEMO_*) with the SBP order (-DSBP_ORDER=2|4|6) as the
// reference's `make <demo> order=N` does (code/Makefile:20-25,45-61).
//
// Command line, problem constants, initial data, analytic solutions and the
// printed lines follow the reference drivers:
//   argv        code/src/util/io_util.cpp:54-154  (1D: N Tend CFL use_custom_sc,
//                                                  2D: Nx Ny Tend CFL use_custom_sc)
//   adv_1D      demo/advection/advection_1D_sim.cpp:55-70,137-157,171-188
//   adv_2D      demo/advection/advection_2D_sim.cpp:61-75,93-97,153-171,188-209
//   wave        demo/wave/wave_eq_sim.cpp:81-91,109-116,173-194,211-233
//   wave_hom    demo/wave_hom/wave_eq_hom_sim.cpp (same lines)
//   reflection  demo/reflection/reflection_sim.cpp:56-74,140-157,177-226
// use_custom_sc selects the reference's halo scatter implementation; one GPU
// owns the whole grid, so it is validated and has no effect.  When the
// environment variable SBP_DEMO_DATA_DIR is set, the run does what the
// reference's write_data switch does (e.g. wave_eq_sim.cpp:136,184-195): it
// writes v_init, v and v_error = |v_analytic - v| (vec_util.h:63-70) as PETSc
// binary vectors (io_util.cpp:4-14) and appends the TSV row of
// io_util.cpp:16-36 to data.tsv, all in $SBP_DEMO_DATA_DIR.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sbp_appctx.hpp"

#ifndef SBP_ORDER
#define SBP_ORDER 4
#endif

#if defined(SBP_DEMO_ADV1D)
static const int kPhysics = SBP_ADVECTION_1D;
static const char *kName = "advection_1D";
#elif defined(SBP_DEMO_ADV2D)
static const int kPhysics = SBP_ADVECTION_2D;
static const char *kName = "advection_2D";
#elif defined(SBP_DEMO_WAVE)
static const int kPhysics = SBP_ACOUSTIC_2D;
static const char *kName = "wave";
#elif defined(SBP_DEMO_WAVE_HOM)
static const int kPhysics = SBP_ACOUSTIC_2D;
static const char *kName = "wave_hom";
#elif defined(SBP_DEMO_REFLECTION)
static const int kPhysics = SBP_REFLECTION_1D;
static const char *kName = "reflection";
#else
#error "define one of SBP_DEMO_ADV1D, SBP_DEMO_ADV2D, SBP_DEMO_WAVE, SBP_DEMO_WAVE_HOM, SBP_DEMO_REFLECTION"
#endif

static const bool kIs1D = kPhysics == SBP_ADVECTION_1D || kPhysics == SBP_REFLECTION_1D;
static const double kPi = 3.14159265358979323846264338327950288419716939937510582097494459;

static void usage(const char *exe)
{
  std::printf("------------------------------ USAGE ------------------------------\n");
  if (kIs1D) {
    std::printf("\"%s N Tend CFL use_custom_sc\"\n\n", exe);
    std::printf("N:\t\tnumber of grid points.\n");
  } else {
    std::printf("\"%s Nx Ny Tend CFL use_custom_sc\"\n\n", exe);
    std::printf("Nx:\t\tnumber of grid points in x-direction.\n");
    std::printf("Ny:\t\tnumber of grid points in y-direction.\n");
  }
  std::printf("Tend:\t\tfinal time.\n");
  std::printf("CFL:\t\tCFL number, dt = CFL*min(dx).\n");
  std::printf("use_custom_sc:\t1 - use custom scatter context, 0 - use PETSc scatter context.\n");
  std::printf("------------------------------ Example ------------------------------\n");
  if (kIs1D)
    std::printf("\"%s 101 1 0.1 1\"\n\n", exe);
  else
    std::printf("\"%s 101 101 1 0.1 1\"\n\n", exe);
}

struct Inputs {
  int nx = 0, ny = 1;
  double tend = 0, cfl = 0;
  int use_custom_sc = 0;
};

// Argument checks with the messages of get_inputs (io_util.cpp:54-154).
static bool get_inputs(int argc, char **argv, Inputs &in)
{
  const int expected = kIs1D ? 4 : 5;
  if (argc != expected + 1) {
    std::printf("Error, wrong number of input arguments. Expected %d arguments, got %d.\n", expected, argc - 1);
    usage(argv[0]);
    return false;
  }
  int k = 1;
  in.nx = std::atoi(argv[k++]);
  if (in.nx <= 0) {
    std::printf(kIs1D ? "Error, second argument wrong. Expected N > 0, got %d.\n"
                      : "Error, first argument wrong. Expected Nx > 0, got %d.\n", in.nx);
    usage(argv[0]);
    return false;
  }
  if (!kIs1D) {
    in.ny = std::atoi(argv[k++]);
    if (in.ny <= 0) {
      std::printf("Error, second argument wrong. Expected Ny > 0, got %d.\n", in.ny);
      usage(argv[0]);
      return false;
    }
  }
  in.tend = std::atof(argv[k++]);
  if (in.tend <= 0) {
    std::printf("Error, third argument wrong. Expected Tend > 0, got %f.\n", in.tend);
    usage(argv[0]);
    return false;
  }
  in.cfl = std::atof(argv[k++]);
  if (in.cfl <= 0) {
    std::printf("Error, fourth argument wrong. Expected CFL > 0, got %f.\n", in.cfl);
    usage(argv[0]);
    return false;
  }
  in.use_custom_sc = std::atoi(argv[k++]);
  if (in.use_custom_sc != 0 && in.use_custom_sc != 1) {
    std::printf("Error, sixth argument wrong. Expected use_custom_sc = 1 or 0, got %d.\n", in.use_custom_sc);
    usage(argv[0]);
    return false;
  }
  return true;
}

// Gaussian of radius rstar = 0.1 centred at c (advection_1D_sim.cpp:171-188,
// theta1 of reflection_sim.cpp:177-183).
static double gauss(double x, double c) { return std::exp(-(x - c) * (x - c) / (0.1 * 0.1)); }

// Initial data (t = 0) or analytic solution at time t, AoS [j][i][c], with
// x = xl + i*h as the drivers' initial/analytic functions use.
static void solution(const sbp_problem &p, double t, std::vector<double> &v)
{
  const double h = (p.xr - p.xl) / (p.nx - 1);
  const double hy = kIs1D ? 1.0 : (p.yr - p.yl) / (p.ny - 1);
  const int C = kPhysics == SBP_ACOUSTIC_2D ? 3 : (kPhysics == SBP_REFLECTION_1D ? 2 : 1);
  v.assign((size_t)p.nx * p.ny * C, 0.0);
  for (int j = 0; j < p.ny; j++) {
    const double y = p.yl + j * hy;
    for (int i = 0; i < p.nx; i++) {
      const double x = p.xl + i * h;
      double *q = &v[((size_t)j * p.nx + i) * C];
      if (kPhysics == SBP_ADVECTION_1D) {
        q[0] = gauss(x, p.ax * t);
      } else if (kPhysics == SBP_ADVECTION_2D) {
        const double dx = x - p.ax * t, dy = y - p.ay * t;
        q[0] = std::exp(-(dx * dx + dy * dy) / (0.1 * 0.1));
      } else if (kPhysics == SBP_ACOUSTIC_2D) {
        // standing mode n = 3, m = 4 (wave_eq_sim.cpp:211-233)
        const double n = 3, m = 4, w = std::sqrt(n * n + m * m);
        q[0] = -n * std::cos(n * kPi * x) * std::sin(m * kPi * y) * std::sin(kPi * w * t) / w;
        q[1] = -m * std::sin(n * kPi * x) * std::cos(m * kPi * y) * std::sin(kPi * w * t) / w;
        q[2] = std::sin(kPi * n * x) * std::sin(kPi * m * y) * std::cos(kPi * w * t);
      } else {
        // reflection: theta2 = -theta1, W = xr - xl (reflection_sim.cpp:185-226)
        const double W = p.xr - p.xl;
        if (t == 0) {
          q[0] = -gauss(x, 0) - gauss(x, 0);
          q[1] = -gauss(x, 0) + gauss(x, 0);
        } else {
          q[0] = gauss(x, W - t) + gauss(x, -W + t);
          q[1] = gauss(x, W - t) - gauss(x, -W + t);
        }
      }
    }
  }
}

// PETSc's binary Vec layout (what VecView writes through a binary viewer, the
// natural ordering of a DMDA vector being the AoS layout on one rank):
// big-endian int32 class id 1211214, int32 length, then big-endian doubles;
// plus the "<file>.info" companion holding -vecload_block_size for dof > 1.
static void put_be32(FILE *f, unsigned v)
{
  const unsigned char b[4] = {(unsigned char)(v >> 24), (unsigned char)(v >> 16), (unsigned char)(v >> 8),
                              (unsigned char)v};
  std::fwrite(b, 1, 4, f);
}

static void write_petsc_vec(const char *dir, const char *name, const std::vector<double> &v, int dof)
{
  mkdir(dir, 0755);
  const std::string path = std::string(dir) + "/" + name;
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) {
    std::printf("File '%s' failed to open.\n", path.c_str());
    return;
  }
  put_be32(f, 1211214u);
  put_be32(f, (unsigned)v.size());
  for (double x : v) {
    unsigned long long u;
    std::memcpy(&u, &x, 8);
    put_be32(f, (unsigned)(u >> 32));
    put_be32(f, (unsigned)u);
  }
  std::fclose(f);
  if (dof > 1) {
    FILE *g = std::fopen((path + ".info").c_str(), "w");
    if (g) {
      std::fprintf(g, "-vecload_block_size %d\n", dof);
      std::fclose(g);
    }
  }
}

static void append_tsv(double dt, const Inputs &in, double elapsed, double l2, double maxerr)
{
  const char *dir = std::getenv("SBP_DEMO_DATA_DIR");
  mkdir(dir, 0755);
  const std::string path = std::string(dir) + "/data.tsv";
  FILE *f = std::fopen(path.c_str(), "a");
  if (!f) {
    std::printf("File '%s' failed to open.\n", path.c_str());
    return;
  }
  std::fseek(f, 0, SEEK_END);
  if (!std::ftell(f)) std::fprintf(f, "Size\tNx\tNy\tdt\tTend\telapsed_time\tl2-error\tH-error\tmax-error\n");
  std::fprintf(f, "%d\t%d\t%d\t%e\t%f\t%f\t%e\t%e\n", 1, in.nx, kIs1D ? -1 : in.ny, dt, in.tend, elapsed, l2, maxerr);
  std::fclose(f);
}

int main(int argc, char **argv)
{
  Inputs in;
  if (!get_inputs(argc, argv, in)) return -1;

  sbp_problem p{};
  p.physics = kPhysics;
  p.order = SBP_ORDER;
  p.nx = in.nx;
  p.ny = kIs1D ? 1 : in.ny;
  p.xl = -1;
  p.xr = 1;
  p.yl = -1;
  p.yr = 1;
  p.material = SBP_MAT_UNIT;
  p.bc = SBP_BC_DEFAULT;
  p.precision = SBP_FP64;
#if defined(SBP_DEMO_ADV1D)
  p.ax = 1.5;  // a(i) = 1.5 (advection_1D_sim.cpp:70)
#elif defined(SBP_DEMO_ADV2D)
  p.ax = 1.5;  // a, b (advection_2D_sim.cpp:74-75)
  p.ay = -1.0;
#elif defined(SBP_DEMO_WAVE)
  p.material = SBP_MAT_REF_RHO;
#endif

  const double hix = (p.nx - 1) / (p.xr - p.xl);
  const double hiy = kIs1D ? hix : (p.ny - 1) / (p.yr - p.yl);
  const double dt = kPhysics == SBP_REFLECTION_1D ? in.cfl * (1.0 / hix) : in.cfl / std::min(hix, hiy);

#if defined(SBP_DEMO_REFLECTION)
  const double L = p.xr - p.xl;
  if (in.tend < L - L / 4 || in.tend > L + L / 4)
    std::printf("--- Warning ---\nTend = %f is outside range of known solution T = [%f,%f]\n", in.tend, L - L / 4,
                L + L / 4);
#endif

  try {
    sbp::AppCtx ctx(p);
    if (!kIs1D) {
      // One GPU owns the whole grid: a 1 x 1 topology.
      std::printf("Processor topology dimensions: [%d,%d]\n", 1, 1);
      std::printf("--- Warning ---\nOne dimensional topology\n");
    }
    std::vector<double> v, v_analytic;
    solution(p, 0.0, v);
    const char *data_dir = std::getenv("SBP_DEMO_DATA_DIR");
    const bool write_data = data_dir && *data_dir;
    if (write_data) write_petsc_vec(data_dir, "v_init", v, ctx.dofs());

    const auto t0 = std::chrono::steady_clock::now();
    sbp::check(sbp::ts_rk4(in.tend, dt, v.data(), ctx));
    const double elapsed = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::printf("Elapsed time: %f seconds\n", elapsed);

    solution(p, in.tend, v_analytic);
    double l2 = 0, maxerr = 0;
    sbp::check(sbp::error_norms(ctx, v.data(), v_analytic.data(), &l2, &maxerr));
    std::printf("The l2-error is: %g, and the maximum error is %g\n", l2, maxerr);
    if (write_data) {
      write_petsc_vec(data_dir, "v", v, ctx.dofs());
      std::vector<double> err(v.size());
      for (size_t k = 0; k < v.size(); k++) err[k] = std::fabs(v_analytic[k] - v[k]);
      write_petsc_vec(data_dir, "v_error", err, ctx.dofs());
      append_tsv(dt, in, elapsed, l2, maxerr);
    }
  } catch (const std::exception &e) {
    std::printf("%s: %s\n", kName, e.what());
    return 1;
  }
  return 0;
}
