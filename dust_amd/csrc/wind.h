// variable_wind (src/aeroel/mod_wind.f90:66-121): the uniform wind plus the
// optional gust, evaluated at a point and a time.  Host and device.
#pragma once
#include <cmath>

#include "common.h"

namespace dust {

// The gust parameters with the directions normalised once (mod_wind.f90:
// 85-89 normalise them on every call; the quotient is the same).
struct WindDev {
  double u_inf[3];
  int type;  // DUST_GUST_NONE / ACM / LINEAR
  double origin[3], fd[3];
  double speed, u_des, gradient, start;
};

inline WindDev wind_dev(const dust_wind_params &w) {
  WindDev d{};
  for (int e = 0; e < 3; ++e) d.u_inf[e] = w.u_inf[e];
  d.type = w.gust_type;
  const double n = std::sqrt(w.front_direction[0] * w.front_direction[0] +
                             w.front_direction[1] * w.front_direction[1] +
                             w.front_direction[2] * w.front_direction[2]);
  for (int e = 0; e < 3; ++e) {
    d.origin[e] = w.origin[e];
    d.fd[e] = n > 0.0 ? w.front_direction[e] / n : 0.0;
  }
  d.speed = w.front_speed;
  d.u_des = w.u_des;
  d.gradient = w.gradient;
  d.start = w.start_time;
  return d;
}

// wind = u_inf; ACM ("1 - cos") gust: s = -(pos - (origin + speed fd
// (time - start))) . fd, the penetration behind the moving front; for
// 0 <= s < 2 gradient every component gains u_des / 2 (1 - cos(pi s /
// gradient)) (the reference adds the scalar to the vector, 108-110).
// "linear" (its test form, 112-124): wind_z += 0.1 pos_x.
__host__ __device__ inline void variable_wind(const WindDev &W, const double pos[3], double time,
                                              double wind[3]) {
#pragma clang fp contract(off)
  wind[0] = W.u_inf[0];
  wind[1] = W.u_inf[1];
  wind[2] = W.u_inf[2];
  if (W.type == DUST_GUST_ACM) {
    const double dt = time - W.start;
    double s = 0.0;
    for (int e = 0; e < 3; ++e) s += (pos[e] - (W.origin[e] + W.speed * W.fd[e] * dt)) * W.fd[e];
    s = -s;
    if (s >= 0.0 && s < 2.0 * W.gradient) {
      const double g = W.u_des / 2.0 * (1.0 - cos(3.141592653589793 * s / W.gradient));
      wind[0] += g;
      wind[1] += g;
      wind[2] += g;
    }
  } else if (W.type == DUST_GUST_LINEAR) {
    wind[2] += pos[0] * 0.1;
  }
}

}  // namespace dust
