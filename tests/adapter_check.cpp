// Compiled and run by tests/test_abi.py on the CPU: exercises the C++ adapter
// (include/mpdg_heat_operator.hpp) error paths that need no GPU.
#include <mpdg_heat_operator.hpp>

#include <cstdio>

int main()
{
  mpdg_heat_desc d{};
  d.n_cells[0] = d.n_cells[1] = d.n_cells[2] = 4;
  d.box_hi[0] = d.box_hi[1] = d.box_hi[2] = 1.0;
  d.degree = 9; // unsupported degree: rejected before any device call
  try
    {
      mpdg::HeatOperator op(d);
      std::printf("no exception\n");
      return 1;
    }
  catch (const mpdg::Error &e)
    {
      if (e.status != MPDG_ERR_UNSUPPORTED)
        return 2;
    }
  mpdg_solver_params p;
  mpdg_solver_params_default(&p);
  if (p.max_iterations != 10000 || p.rel_tolerance != 1e-12)
    return 3;
  std::printf("adapter ok\n");
  return 0;
}
