// Initial fields of the periodic-box workloads, evaluated at the cell centroids
// the way an INIT expression or an initialHook sets them
// (StaticSolver.hpp:1402-1470, charles_ig.cpp:209-222). Host C++, threaded.
//
// Taylor-Green vortex on [0, 2 pi]^3:
//   u = U0 sin x cos y cos z, v = -U0 cos x sin y cos z, w = 0,
//   p = p0 + rho0 U0^2/16 (cos 2x + cos 2y)(cos 2z + 2), rho = p/(R T0),
//   rhoE = p/(gamma-1) + rho |u|^2/2   (IdealGasSolver.hpp:282-285)
// with U0 = rho0 = T0 = 1, p0 = 1/(gamma Ma^2).
//
// Solenoidal random field with E(k) ~ k^4 exp(-2 (k/k0)^2): a sum of Fourier modes
// over integer wave vectors |k| <= kmax (half space), each with an amplitude
// sqrt(E(|k|) / (2 pi |k|^2)), a random direction orthogonal to k and a random
// phase (seeded), the mean removed and scaled so u_rms/c0 = Mt with c0 = 1;
// uniform rho = 1, p = 1/gamma.
#include <cmath>
#include <complex>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "mesh.hpp"

namespace cgmesh {

void ic_tgv(int ncv, const double* x_cv, double Ma, double gamma, double* rho, double* u, double* rhoE) {
  const double c0 = 1.0 / Ma;
  const double p0 = c0 * c0 / gamma;
  const double R = p0;
  for (int i = 0; i < ncv; ++i) {
    const double x = x_cv[3 * i], y = x_cv[3 * i + 1], z = x_cv[3 * i + 2];
    const double u0 = std::sin(x) * std::cos(y) * std::cos(z);
    const double u1 = -std::cos(x) * std::sin(y) * std::cos(z);
    const double p = p0 + (1.0 / 16.0) * (std::cos(2 * x) + std::cos(2 * y)) * (std::cos(2 * z) + 2.0);
    rho[i] = p / (R * 1.0);
    u[3 * i] = u0;
    u[3 * i + 1] = u1;
    u[3 * i + 2] = 0.0;
    rhoE[i] = p / (gamma - 1.0) + 0.5 * rho[i] * (u0 * u0 + u1 * u1 + 0.0 * 0.0);
  }
}

std::string ic_k4(int ncv, const double* x_cv, double k0, double Mt, double gamma, uint64_t seed, int kmax,
                  double* rho, double* u, double* rhoE, int nthreads) {
  if (kmax <= 0) kmax = int(std::ceil(2.5 * k0));
  if (kmax > 64) return "kmax too large";
  struct Mode { int k[3]; double a[3]; double phase; };
  std::vector<Mode> modes;
  uint64_t ctr = 0;
  auto normal = [&]() {
    const double u1 = hash_uniform(seed, ctr++), u2 = hash_uniform(seed, ctr++);
    return std::sqrt(-2.0 * std::log(u1 > 0 ? u1 : 1e-300)) * std::cos(2.0 * M_PI * u2);
  };
  for (int a = -kmax; a <= kmax; ++a)
    for (int b = -kmax; b <= kmax; ++b)
      for (int c = 0; c <= kmax; ++c) {
        if (c == 0 && (b < 0 || (b == 0 && a <= 0))) continue;
        if (a * a + b * b + c * c > kmax * kmax) continue;
        Mode m;
        m.k[0] = a; m.k[1] = b; m.k[2] = c;
        const double km = std::sqrt(double(a * a + b * b + c * c));
        const double E = std::pow(km, 4) * std::exp(-2.0 * (km / k0) * (km / k0));
        const double amp = std::sqrt(E / (2.0 * M_PI * km * km));
        double r[3] = {normal(), normal(), normal()};
        double d[3] = {b * r[2] - c * r[1], c * r[0] - a * r[2], a * r[1] - b * r[0]};
        double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (dn == 0.0) { d[0] = 1; d[1] = 0; d[2] = 0; dn = 1; }
        for (int q = 0; q < 3; ++q) m.a[q] = amp * d[q] / dn;
        m.phase = 2.0 * M_PI * hash_uniform(seed, ctr++);
        modes.push_back(m);
      }
  // u(x) = sum_m a_m cos(k_m . x + phase_m), via per-cell tables of exp(i k x_d)
  std::vector<std::thread> th;
  if (nthreads < 1) nthreads = 1;
  for (int t = 0; t < nthreads; ++t)
    th.emplace_back([&, t] {
      const int K = 2 * kmax + 1;
      std::vector<std::complex<double>> ex(K), ey(K), ez(K);
      for (int i = t; i < ncv; i += nthreads) {
        const double* x = &x_cv[3 * i];
        for (int q = -kmax; q <= kmax; ++q) {
          ex[q + kmax] = std::polar(1.0, q * x[0]);
          ey[q + kmax] = std::polar(1.0, q * x[1]);
          ez[q + kmax] = std::polar(1.0, q * x[2]);
        }
        double v[3] = {0, 0, 0};
        for (const Mode& m : modes) {
          const std::complex<double> e = ex[m.k[0] + kmax] * ey[m.k[1] + kmax] * ez[m.k[2] + kmax] *
                                         std::polar(1.0, m.phase);
          for (int q = 0; q < 3; ++q) v[q] += m.a[q] * e.real();
        }
        for (int q = 0; q < 3; ++q) u[3 * i + q] = v[q];
      }
    });
  for (auto& x : th) x.join();
  double mean[3] = {0, 0, 0};
  for (int i = 0; i < ncv; ++i)
    for (int q = 0; q < 3; ++q) mean[q] += u[3 * i + q];
  for (int q = 0; q < 3; ++q) mean[q] /= ncv;
  double ss = 0.0;
  for (int i = 0; i < ncv; ++i)
    for (int q = 0; q < 3; ++q) {
      u[3 * i + q] -= mean[q];
      ss += u[3 * i + q] * u[3 * i + q];
    }
  const double urms = std::sqrt(ss / (3.0 * ncv));
  if (!(urms > 0.0)) return "degenerate velocity field";
  const double scale = Mt * 1.0 / urms;
  const double p0 = 1.0 / gamma;
  for (int i = 0; i < ncv; ++i) {
    double usq = 0.0;
    for (int q = 0; q < 3; ++q) {
      u[3 * i + q] *= scale;
      usq += u[3 * i + q] * u[3 * i + q];
    }
    rho[i] = 1.0;
    rhoE[i] = p0 / (gamma - 1.0) + 0.5 * usq;
  }
  return "";
}

}  // namespace cgmesh
