// Run by tests/test_gpu_adapter.py on the GPU: the header-only C++ adapter
// (include/mpdg_heat_operator.hpp) driving a device operator the way host code
// of the reference drives HeatDiffuseMultiPhaseOperator:
//   op.vmult(dst, src); JacobiPreconditioner::update(); LinearSolver::solve(...)
// Usage: adapter_gpu n degree src.bin rhs.bin out.bin
//   the operator is the unit cube with n^3 cells, rho_cp = 1, dt = 1e-3,
//   k = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z), Dirichlet on all faces;
//   out.bin receives vmult(src), then the solution of CG + Jacobi to 1e-10 with
//   a caller-updated preconditioner, then the one of the solve that updates it
//   itself; stdout prints both iteration counts.
#include <mpdg_heat_operator.hpp>

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

namespace
{
  bool read_file(const char *path, std::vector<double> &v)
  {
    FILE *f = std::fopen(path, "rb");
    if (!f)
      return false;
    const size_t got = std::fread(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
    return got == v.size();
  }

  double *device_vector(size_t n)
  {
    void *p = nullptr;
    if (hipMalloc(&p, n * sizeof(double)) != hipSuccess)
      {
        std::fprintf(stderr, "hipMalloc failed\n");
        std::exit(3);
      }
    return static_cast<double *>(p);
  }
} // namespace

int main(int argc, char **argv)
{
  if (argc != 6)
    {
      std::fprintf(stderr, "usage: adapter_gpu n degree src.bin rhs.bin out.bin\n");
      return 1;
    }
  const int n = std::atoi(argv[1]);
  mpdg_heat_desc d{};
  d.n_cells[0] = d.n_cells[1] = d.n_cells[2] = n;
  d.degree = std::atoi(argv[2]);
  d.box_hi[0] = d.box_hi[1] = d.box_hi[2] = 1.0;
  d.precision = MPDG_FP64;
  d.material = MPDG_MATERIAL_GAS;
  d.gas.density = d.gas.specific_heat_capacity = d.gas.thermal_conductivity = 1.0;
  d.conductivity_field = MPDG_CONDUCTIVITY_SEPARABLE_SINE;
  d.conductivity_amplitude = 0.5;
  d.dirichlet_faces = 0x3F;
  try
    {
      mpdg::HeatOperator op(d);
      op.reset_time_increment(1e-3);
      const size_t N = op.n_dofs();
      std::vector<double> src(N), rhs(N), out(3 * N);
      if (!read_file(argv[3], src) || !read_file(argv[4], rhs))
        {
          std::fprintf(stderr, "could not read the input vectors\n");
          return 2;
        }
      mpdg::DeviceVector s{device_vector(N), N}, y{device_vector(N), N}, b{device_vector(N), N},
        x{device_vector(N), N}, diag{device_vector(N), N};
      if (hipMemcpy(s.data, src.data(), N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(b.data, rhs.data(), N * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return 3;
      op.vmult(y, s);

      mpdg::JacobiPreconditioner prec(op, diag);
      prec.update();
      mpdg::LinearSolverData data;
      data.solver_type = MPDG_SOLVER_CG;
      data.preconditioner_type = MPDG_PRECONDITIONER_DIAGONAL;
      data.rel_tolerance = 1e-10;
      if (hipMemset(x.data, 0, N * sizeof(double)) != hipSuccess)
        return 3;
      const int it1 = mpdg::LinearSolver::solve(op, x, b, data, prec);
      if (hipMemcpy(out.data() + N, x.data, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        return 3;
      if (hipMemset(x.data, 0, N * sizeof(double)) != hipSuccess)
        return 3;
      const int it2 = mpdg::LinearSolver::solve(op, x, b, data);
      if (hipMemcpy(out.data(), y.data, N * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess ||
          hipMemcpy(out.data() + 2 * N, x.data, N * sizeof(double), hipMemcpyDeviceToHost) !=
            hipSuccess)
        return 3;
      FILE *f = std::fopen(argv[5], "wb");
      if (!f || std::fwrite(out.data(), sizeof(double), out.size(), f) != out.size())
        return 4;
      std::fclose(f);
      for (double *p : {s.data, y.data, b.data, x.data, diag.data})
        (void)hipFree(p);
      std::printf("iterations %d %d\n", it1, it2);
    }
  catch (const mpdg::Error &e)
    {
      std::fprintf(stderr, "adapter error %d: %s\n", (int)e.status, e.what());
      return 5;
    }
  return 0;
}
