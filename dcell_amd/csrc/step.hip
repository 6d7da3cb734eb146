// step.hip -- one dense level-set update: SL step phi0 -> phi followed by
// the 19-point (3D) / 9-point (2D) curvature of the new phi.
#include "dcl_host.hpp"

namespace dcl {
int sl_launch(const void* phi0, void* phi, const dcl_window* win, const dcl_velocity* vel,
              const dcl_icoor* band, int64_t band_len, void* band_values, double dt,
              int precision, int interp, int backtrace, hipStream_t s);
int check_sl_args(const void* phi0, void* phi, const dcl_window* win, const dcl_velocity* vel,
                  const dcl_icoor* band, int64_t band_len, void* band_values, int precision,
                  int interp, int backtrace);
int curv_launch(const void* phi, void* kappa, const dcl_window* win, const dcl_icoor* band,
                int64_t n, void* bv, int precision, int stencil, hipStream_t s);
}  // namespace dcl

using namespace dcl;

extern "C" int dcl_levelset_step(const void* phi0, void* phi, void* kappa, const dcl_window* win,
                                 const dcl_velocity* vel, double dt, int precision, int interp,
                                 int backtrace, void* stream) {
  DCL_TRY(check_sl_args(phi0, phi, win, vel, nullptr, 0, nullptr, precision, interp, backtrace));
  if (!kappa) return set_error(DCL_ERR_ARG_NULL, "kappa is NULL");
  if (kappa == phi || kappa == phi0)
    return set_error(DCL_ERR_ARG_WRONG, "kappa must not alias phi or phi0");
  hipStream_t s = as_stream(stream);
  DCL_TRY(sl_launch(phi0, phi, win, vel, nullptr, 0, nullptr, dt, precision, interp, backtrace, s));
  return curv_launch(phi, kappa, win, nullptr, 0, nullptr, precision, DCL_CURV_STD, s);
}
