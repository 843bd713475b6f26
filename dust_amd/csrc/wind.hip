// variable_wind at many points (src/aeroel/mod_wind.f90:66-121), for the
// callers that need the wind at panel centres, wake points or particles.
#include "common.h"
#include "wind.h"

namespace dust {
namespace {

__global__ void k_wind(int n, const double *__restrict__ pos, WindDev W, double time,
                       double *wind) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p[3] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
  double w[3];
  variable_wind(W, p, time, w);
  wind[3 * i] = w[0];
  wind[3 * i + 1] = w[1];
  wind[3 * i + 2] = w[2];
}

int wind_check(int n, const double *pos, const dust_wind_params *w, const double *wind) {
  DUST_CHECK(n >= 0, DUST_E_ARG, "negative point count");
  DUST_CHECK(w, DUST_E_ARG, "null wind parameters");
  DUST_CHECK(w->gust_type >= DUST_GUST_NONE && w->gust_type <= DUST_GUST_LINEAR, DUST_E_ARG,
             "unknown gust type %d", w->gust_type);
  DUST_CHECK(n == 0 || (pos && wind), DUST_E_ARG, "null point or wind array");
  return 0;
}

}  // namespace
}  // namespace dust

using namespace dust;

extern "C" {

int dust_variable_wind_dev(int n, const double *pos, const dust_wind_params *w, double time,
                           double *wind, void *stream) {
  DUST_TRY(wind_check(n, pos, w, wind));
  if (n == 0) return 0;
  k_wind<<<ceil_div(n, 256), 256, 0, (hipStream_t)stream>>>(n, pos, wind_dev(*w), time, wind);
  DUST_HIP(hipGetLastError());
  return 0;
}

int dust_variable_wind(int n, const double *pos, const dust_wind_params *w, double time,
                       double *wind) {
  DUST_TRY(wind_check(n, pos, w, wind));
  if (n == 0) return 0;
  TmpDev<double> dp, dw;
  DUST_TRY(dp.alloc((size_t)3 * n));
  DUST_TRY(dw.alloc((size_t)3 * n));
  DUST_HIP(hipMemcpy(dp.p, pos, sizeof(double) * 3 * n, hipMemcpyHostToDevice));
  k_wind<<<ceil_div(n, 256), 256>>>(n, dp.p, wind_dev(*w), time, dw.p);
  DUST_HIP(hipGetLastError());
  DUST_HIP(hipMemcpy(wind, dw.p, sizeof(double) * 3 * n, hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"
