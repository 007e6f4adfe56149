// Host-side initial sampling: the thesis code's ran2 and the two
// Thomas-Fermi samplers.  Native because the production cloud has 8e6 to 8e7
// particles and the generator is sequential by construction.
//
// ran2 (src/random.jl:288-347) is Numerical Recipes' combined L'Ecuyer
// generator with the Bays-Durham shuffle, kept with the thesis's IQ1 = 3668
// and its 32-bit wrap-around, which the reference reproduces on purpose.
// The Int32 products are evaluated in uint32 so that the wrap is defined.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/vlasov_hip.h"

void vh_set_error(const char *fmt, ...);

namespace {

const int32_t IM1 = 2147483563, IM2 = 2147483399, IMM1 = IM1 - 1;
const int32_t IA1 = 40014, IA2 = 40692, IQ2 = 52774, IR1 = 12211, IR2 = 3791;
const int NTAB = 32;
const int32_t NDIV = 1 + (IMM1 - 1) / NTAB;

inline int32_t wrap_mul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
inline int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
inline int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }

// one Schrage step, overflowing as the Fortran did
inline int32_t ran2_step(int32_t x, int32_t q, int32_t r, int32_t a, int32_t m) {
  int32_t k = x / q;
  x = wrap_sub(wrap_mul(a, wrap_sub(x, wrap_mul(k, q))), wrap_mul(k, r));
  return x < 0 ? wrap_add(x, m) : x;
}

struct Ran2 {
  int32_t idum, idum2, iy, iq1;
  int32_t iv[NTAB];
  float am, rnmx;

  Ran2(int seed, bool consistent) {
    iq1 = consistent ? 53668 : 3668;
    idum = (int32_t)(-seed > 1 ? -seed : 1);
    idum2 = idum;
    for (int j = 0; j < NTAB; ++j) iv[j] = 0;
    for (int j = NTAB + 8; j >= 1; --j) {
      idum = ran2_step(idum, iq1, IR1, IA1, IM1);
      if (j <= NTAB) iv[j - 1] = idum;
    }
    iy = iv[0];
    am = 1.0f / (float)IM1;
    rnmx = 1.0f - 1.2e-7f;
  }

  float next() {
    idum = ran2_step(idum, iq1, IR1, IA1, IM1);
    idum2 = ran2_step(idum2, IQ2, IR2, IA2, IM2);
    int j = iy / NDIV;  // 0-based
    iy = wrap_sub(iv[j], idum2);
    iv[j] = idum;
    if (iy < 1) iy = wrap_add(iy, IMM1);
    float v = am * (float)iy;
    return v < rnmx ? v : rnmx;
  }
};

// linear interpolation on a regular grid at reduced u (src/initial.jl:71-78)
inline double interp_regular(const double *v, double u, int n) {
  long j = (long)std::floor(u * (n - 1)) + 1;
  if (j > n - 1) j = n - 1;
  const double a = (double)(j - 1) / (n - 1);
  const double b = (double)j / (n - 1);
  return v[j - 1] + (v[j] - v[j - 1]) * (u - a) / (b - a);
}

// point of norm r on the sphere (src/initial.jl:233-238)
inline void on_sphere(double r, double uphi, double umu, double *out) {
  const double mu = 2.0 * umu - 1.0;
  const double s = std::sqrt(1.0 - mu * mu);
  const double phi = 2.0 * M_PI * uphi;
  out[0] = r * std::cos(phi) * s;
  out[1] = r * std::sin(phi) * s;
  out[2] = r * mu;
}

const double FERMI_COEFFICIENT = std::cbrt(3.0 * (M_PI * M_PI));

}  // namespace

extern "C" int vh_ran2_sequence(int seed, int consistent, int64_t count, float *out) {
  if (!out || count < 0) {
    vh_set_error("vh_ran2_sequence: bad argument");
    return VH_EINVAL;
  }
  Ran2 g(seed, consistent != 0);
  for (int64_t i = 0; i < count; ++i) out[i] = g.next();
  return VH_OK;
}

extern "C" int vh_sample_radial(const double *quantiles, int nq, const double *density, int nrho,
                                double rmax, int64_t npart, double weight, int seed, double *pos,
                                double *mom) {
  if (!quantiles || !density || !pos || !mom || nq < 2 || nrho < 2 || npart < 0 || !(rmax > 0)) {
    vh_set_error("vh_sample_radial: bad argument");
    return VH_EINVAL;
  }
  Ran2 g(seed, false);
  for (int64_t i = 0; i < npart; ++i) {
    double x[6];
    for (int k = 0; k < 6; ++k) x[k] = (double)g.next();
    const double r = interp_regular(quantiles, x[0], nq);
    on_sphere(r, x[1], x[2], pos + 3 * i);
    const double rho = interp_regular(density, r / rmax, nrho);
    const double p = std::cbrt(x[3]) * FERMI_COEFFICIENT * std::cbrt(rho);
    on_sphere(p * weight, x[4], x[5], mom + 3 * i);
  }
  return VH_OK;
}

// sample_thomas_fermi(::RadialProfile, n, w; rng) in slices: the generator's
// state carries over from one call to the next, so consecutive slices are
// the particles a single call would have drawn, in order
struct vh_sampler {
  Ran2 g;
  std::vector<double> quantiles, density;
  double rmax, weight;
  vh_sampler(int seed) : g(seed, false) {}
};

extern "C" int vh_sampler_create(const double *quantiles, int nq, const double *density, int nrho,
                                 double rmax, double weight, int seed, vh_sampler **out) {
  if (!quantiles || !density || !out || nq < 2 || nrho < 2 || !(rmax > 0)) {
    vh_set_error("vh_sampler_create: bad argument");
    return VH_EINVAL;
  }
  vh_sampler *s = new vh_sampler(seed);
  s->quantiles.assign(quantiles, quantiles + nq);
  s->density.assign(density, density + nrho);
  s->rmax = rmax;
  s->weight = weight;
  *out = s;
  return VH_OK;
}

extern "C" int vh_sampler_next(vh_sampler *s, int64_t count, double *pos, double *mom) {
  if (!s || !pos || !mom || count < 0) {
    vh_set_error("vh_sampler_next: bad argument");
    return VH_EINVAL;
  }
  const int nq = (int)s->quantiles.size(), nrho = (int)s->density.size();
  for (int64_t i = 0; i < count; ++i) {
    double x[6];
    for (int k = 0; k < 6; ++k) x[k] = (double)s->g.next();
    const double r = interp_regular(s->quantiles.data(), x[0], nq);
    on_sphere(r, x[1], x[2], pos + 3 * i);
    const double rho = interp_regular(s->density.data(), r / s->rmax, nrho);
    const double p = std::cbrt(x[3]) * FERMI_COEFFICIENT * std::cbrt(rho);
    on_sphere(p * s->weight, x[4], x[5], mom + 3 * i);
  }
  return VH_OK;
}

extern "C" int vh_sampler_destroy(vh_sampler *s) {
  delete s;
  return VH_OK;
}

extern "C" int vh_sample_rejection(const double *grid, const double *potential, int ngrid, double rmax,
                                   double pmax, double fermi, int64_t npart, double weight, int seed,
                                   double *pos, double *mom) {
  if (!grid || !potential || !pos || !mom || ngrid < 2 || npart < 0 || !(rmax > 0)) {
    vh_set_error("vh_sample_rejection: bad argument");
    return VH_EINVAL;
  }
  Ran2 g(seed, false);
  const int n = ngrid - 1;
  const double scale = n / rmax;
  for (int64_t i = 0; i < npart; ++i) {
    double x[6], r, p;
    for (;;) {
      for (int k = 0; k < 6; ++k) x[k] = (double)g.next();
      r = rmax * std::cbrt(x[0]);
      p = pmax * std::cbrt(x[3]);
      long j = (long)std::trunc(scale * r - 1e-10);  // 0-based cell
      if (j < 0) j = 0;
      if (j > n - 1) j = n - 1;
      const double v = potential[j] + (potential[j + 1] - potential[j]) * (r - grid[j]) /
                                          (grid[j + 1] - grid[j]);
      if (p * p / 2.0 + v < fermi) break;
    }
    on_sphere(r, x[1], x[2], pos + 3 * i);
    on_sphere(p * weight, x[4], x[5], mom + 3 * i);
  }
  return VH_OK;
}
