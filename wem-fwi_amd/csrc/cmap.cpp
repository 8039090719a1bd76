// Conformal maps of the acquisition geometry (propagator/src/PropParam.h:89-131
// and conformal/src): for every shot, the propagation grid, the gathers that
// carry a model onto it and back, the Jacobian factors applied to the
// slowness, and the mapped device positions.  Host only.  The arithmetic is
// std::complex<float> in the order of the reference's expressions, because the
// truncations of mapped coordinates to grid indices decide which samples are
// used.
#include <algorithm>
#include <cmath>
#include <complex>
#include <string>
#include <vector>

#include "../../include/wemfwi.h"

namespace wemfwi {

int fail(int code, const std::string& msg);  // host.cpp: records the message, returns code

namespace {

using cf = std::complex<float>;

// Regular grid: axis 1 = x (fast), axis 2 = z.
struct Grid {
  int n1 = 0, n2 = 0;
  float o1 = 0.f, d1 = 1.f, o2 = 0.f, d2 = 1.f;
  size_t size() const { return (size_t)n1 * n2; }
  // coordinate z + i x of sample i (ConformalMap.h:98-109)
  cf coord(size_t i) const {
    const int i1 = (int)(i % n1), i2 = (int)(i / n1);
    return cf(o2 + i2 * d2, o1 + i1 * d1);
  }
  // truncated grid indices of a coordinate; false outside the grid
  bool cell(cf c, int& iz, int& ix) const {
    iz = (int)((c.real() - o2) / d2);
    ix = (int)((c.imag() - o1) / d1);
    return iz >= 0 && iz < n2 && ix >= 0 && ix < n1;
  }
};

// One map g: Z -> U.  zmap[i]: the Z coordinate of output sample i (z =
// g^-1(u)); umap[i]: the U coordinate of input sample i; jac[i]: Jacobian
// factor of output sample i.
struct Map {
  Grid in, out;
  std::vector<cf> zmap, umap;
  std::vector<float> jac;
  void alloc() {
    zmap.resize(in.size());
    umap.resize(in.size());
    jac.resize(in.size());
  }
};

const float kPi = 4 * std::atan(1.f);

// Products and quotients with the real scale a of ShiftAndScale, written out
// per component: PropParam builds a as a real number (PropParam.h:100-103),
// and then the complex operations reduce to these exactly (a complex quotient
// by (a, 0) is the correctly rounded per-component quotient in the C
// library's division; inline compiler expansions of complex division may
// round differently).
cf scale(float a, cf z) { return cf(a * z.real(), a * z.imag()); }
cf unscale(cf z, float a) { return cf(z.real() / a, z.imag() / a); }

// f(z) = a (z - b) with real a  (ShiftAndScale.h:15-55)
Map shift_scale(const Grid& g, float a, cf b) {
  Map m;
  m.in = g;
  cf lo(g.o2, g.o1);
  cf hi(g.o2 + (g.n2 - 1) * g.d2, g.o1 + (g.n1 - 1) * g.d1);
  lo = scale(a, lo - b);
  hi = scale(a, hi - b);
  m.out.n1 = g.n1;
  m.out.n2 = g.n2;
  m.out.o1 = lo.imag();
  m.out.d1 = (hi - lo).imag() / (g.n1 - 1);
  m.out.o2 = lo.real();
  m.out.d2 = (hi - lo).real() / (g.n2 - 1);
  m.alloc();
  const float j = std::fabs(1.f / a);
  for (size_t i = 0; i < g.size(); ++i) {
    m.zmap[i] = unscale(m.out.coord(i), a) + b;
    m.umap[i] = scale(a, g.coord(i) - b);
    m.jac[i] = j;
  }
  return m;
}

// u = log z on the annulus e^eps <= |z| <= 1 (ComplexLog.h:14-66): output
// axis 1 is the angle 0..2 pi, axis 2 the log radius log(eps)..0.
Map complex_log(const Grid& g, float eps_circle) {
  Map m;
  m.in = g;
  const float e = std::log(eps_circle);
  m.out.n1 = g.n1;
  m.out.n2 = g.n2;
  m.out.o1 = 0.f;
  m.out.d1 = 2 * kPi / (g.n1 - 1);
  m.out.o2 = e;
  m.out.d2 = -e / (g.n2 - 1);
  m.alloc();
  const float rmin = std::exp(e);
  const cf ip(0.f, kPi);
  for (size_t i = 0; i < g.size(); ++i) {
    const cf u = m.out.coord(i);
    m.zmap[i] = std::exp(u);
    const cf z = g.coord(i);
    cf v = std::abs(z) < rmin ? cf(e, std::arg(z)) : std::log(z);
    if (v.imag() < 0) v += 2.f * ip;
    m.umap[i] = v;
    m.jac[i] = std::exp(u.real());
  }
  return m;
}

// rotation by -rad about the origin (Rotation.h:14-99), with the clamping of
// out-of-grid preimages to the grid edges as written (including its member
// extents starting at 0).
Map rotation(const Grid& g, float rad) {
  Map m;
  m.in = g;
  const cf I(0.f, 1.f);
  const cf rot = std::exp(-I * rad), irot = std::exp(I * rad);
  const cf v[4] = {cf(g.o2, g.o1), cf(g.o2, g.o1 + (g.n1 - 1) * g.d1), cf(g.o2 + (g.n2 - 1) * g.d2, g.o1),
                   cf(g.o2 + (g.n2 - 1) * g.d2, g.o1 + (g.n1 - 1) * g.d1)};
  float o1 = 0.f, o2 = 0.f, max1 = 0.f, max2 = 0.f;
  for (int k = 0; k < 4; ++k) {
    const cf u = v[k] * rot;
    o1 = std::min(o1, u.imag());
    o2 = std::min(o2, u.real());
    max1 = std::max(max1, u.imag());
    max2 = std::max(max2, u.real());
  }
  m.out.n1 = g.n1;
  m.out.n2 = g.n2;
  m.out.o1 = o1;
  m.out.d1 = (max1 - o1) / (g.n1 - 1);
  m.out.o2 = o2;
  m.out.d2 = (max2 - o2) / (g.n2 - 1);
  m.alloc();
  const float m1 = (g.n1 - 1) * g.d1 + g.o1;
  const float m2 = (g.n2 - 1) * g.d2 + g.o2;
  const float t = std::tan(rad);
  for (size_t i = 0; i < g.size(); ++i) {
    cf z = m.out.coord(i) * irot;
    const int i1 = (int)((z.imag() - g.o1) / g.d1);
    const int i2 = (int)((z.real() - g.o2) / g.d2);
    float x;
    if (i1 < 0) {
      x = -(z.imag() - g.o1) / t;
      x = std::min(z.real() + x, m2);
      z = cf(x, g.o1);
    }
    if (i2 < 0) {
      x = -(z.real() - g.o2) * t;
      x = std::min(z.imag() + x, m1);
      z = cf(g.o2, x);
    }
    if (i1 >= g.n1) {
      x = (z.imag() - m1) / t;
      x = std::max(z.real() - x, g.o1);
      z = cf(x, m1);
    }
    if (i2 >= g.n2) {
      x = (z.real() - m2) * t;
      x = std::min(z.imag() - x, g.o1);
      z = cf(m2, x);
    }
    m.zmap[i] = z;
    m.umap[i] = g.coord(i) * rot;
    m.jac[i] = 1.f;
  }
  return m;
}

// Gather of ConformalMap::forward (ConformalMap.cpp:53-72): out[i] = in[cell of
// zmap[i]] where that cell is inside the input grid, -1 (untouched) elsewhere.
std::vector<int> forward_gather(const Map& m) {
  std::vector<int> idx(m.in.size(), -1);
  for (size_t i = 0; i < m.in.size(); ++i) {
    int iz, ix;
    if (m.in.cell(m.zmap[i], iz, ix)) idx[i] = iz * m.in.n1 + ix;
  }
  return idx;
}

// Gather of ComplexLog::forward(const&, &) (ComplexLog.cpp:77-109): a sample
// whose preimage is outside the input keeps the value of the last sample of
// its column that had one (the row buffer is not cleared), or zero.
std::vector<int> log_forward_gather(const Map& m) {
  std::vector<int> idx(m.in.size(), -1), col(m.in.n1, -1);
  for (int i2 = 0; i2 < m.in.n2; ++i2)
    for (int i1 = 0; i1 < m.in.n1; ++i1) {
      const size_t i = (size_t)i2 * m.in.n1 + i1;
      int iz, ix;
      if (m.in.cell(m.zmap[i], iz, ix)) col[i1] = iz * m.in.n1 + ix;
      idx[i] = col[i1];
    }
  return idx;
}

// Gather of ConformalMap::inverse (ConformalMap.cpp:74-87): in[i] = out[cell of
// umap[i]] inside the output grid, -1 (untouched) elsewhere.
std::vector<int> inverse_gather(const Map& m) {
  std::vector<int> idx(m.in.size(), -1);
  for (size_t i = 0; i < m.in.size(); ++i) {
    int iz, ix;
    if (m.out.cell(m.umap[i], iz, ix)) idx[i] = iz * m.out.n1 + ix;
  }
  return idx;
}

// Gather of ComplexLog::inverse (ComplexLog.cpp:111-130): only samples with
// e^eps < |z| < 0.99 are written (no bounds test, as in the reference).
std::vector<int> log_inverse_gather(const Map& m, float eps_circle) {
  std::vector<int> idx(m.in.size(), -1);
  const float rmin = std::exp(std::log(eps_circle));
  for (size_t i = 0; i < m.in.size(); ++i) {
    const float amp = std::abs(m.in.coord(i));
    if (amp > rmin && amp < .99f) {
      int iz, ix;
      m.out.cell(m.umap[i], iz, ix);
      if (iz < 0 || iz >= m.out.n2 || ix < 0 || ix >= m.out.n1) return {};  // would read outside the grid
      idx[i] = iz * m.out.n1 + ix;
    }
  }
  return idx;
}

// Device position of ConformalMap::forward(int&, int&) (ConformalMap.h:34-48).
void map_point(const Map& m, int iz, int ix, int& oz, int& ox) {
  const cf c = m.umap[(size_t)ix + (size_t)iz * m.in.n1];
  oz = (int)((c.real() - m.out.o2) / m.out.d2);
  ox = (int)((c.imag() - m.out.o1) / m.out.d1);
}

}  // namespace
}  // namespace wemfwi

using namespace wemfwi;

extern "C" {

int wemfwi_cmap_plan(int kind, int nz, int nx, float oz, float ox, float dz, float dx, int ns, const int* src_iz,
                     const int* src_ix, int nr, const int* rec_iz, const int* rec_ix, float eps_log,
                     const float* angles, float* axes, int* fwd, int* inv, float* jac, int* nsrc, int* siz,
                     int* six, int* riz, int* rix) {
  if (kind < 0 || kind > 2) return fail(WEMFWI_ERR_ARG, "cmap: kind must be 0, 1 or 2");
  if (nz < 2 || nx < 2 || ns < 1 || nr < 1 || !src_iz || !src_ix || !rec_iz || !rec_ix)
    return fail(WEMFWI_ERR_ARG, "cmap: invalid geometry");
  if (!(dz > 0.f) || !(dx > 0.f)) return fail(WEMFWI_ERR_ARG, "cmap: invalid grid spacing");
  if (!axes || !fwd || !inv || !jac || !nsrc || !siz || !six || !riz || !rix)
    return fail(WEMFWI_ERR_ARG, "cmap: NULL output");
  if (kind == 1 && !(eps_log > 0.f && eps_log < 1.f)) return fail(WEMFWI_ERR_ARG, "cmap: eps_log must be in (0, 1)");
  if (kind == 1 && nr == 1)
    return fail(WEMFWI_ERR_UNSUPPORTED, "cmap 1: a single receiver becomes a line of nx receivers (ComplexLog.cpp:"
                                        "132-165); not supported");
  if (kind == 2 && !angles) return fail(WEMFWI_ERR_ARG, "cmap 2: angles required");
  const Grid g{nx, nz, ox, dx, oz, dz};
  const size_t n = g.size();
  for (int i = 0; i < nr; ++i)
    if (rec_iz[i] < 0 || rec_iz[i] >= nz || rec_ix[i] < 0 || rec_ix[i] >= nx)
      return fail(WEMFWI_ERR_ARG, "cmap: receiver outside the model");
  float a = 0.f;  // PropParam.h:91: kept across shots
  for (int s = 0; s < ns; ++s) {
    if (src_iz[s] < 0 || src_iz[s] >= nz || src_ix[s] < 0 || src_ix[s] >= nx)
      return fail(WEMFWI_ERR_ARG, "cmap: source outside the model");
    float* ax = axes + 4 * (size_t)s;
    int* f = fwd + (size_t)s * n;
    int* v = inv + (size_t)s * n;
    float* j1 = jac + (size_t)s * 2 * n;
    float* j2 = j1 + n;
    int* sz = siz + (size_t)s * nx;
    int* sx = six + (size_t)s * nx;
    int* rz = riz + (size_t)s * nr;
    int* rx = rix + (size_t)s * nr;
    Grid out = g;
    if (kind == 0) {
      for (size_t i = 0; i < n; ++i) {
        f[i] = (int)i;
        v[i] = (int)i;
        j1[i] = 1.f;
        j2[i] = 1.f;
      }
      nsrc[s] = 1;
      sz[0] = src_iz[s];
      sx[0] = src_ix[s];
      for (int i = 0; i < nr; ++i) {
        rz[i] = rec_iz[i];
        rx[i] = rec_ix[i];
      }
    } else if (kind == 1) {
      // PropParam.h:90-110: shift the source to the origin, scale the farthest
      // corner to |z| = 1, then take the complex log
      const cf b(g.o2 + src_iz[s] * g.d2, g.o1 + src_ix[s] * g.d1);
      for (int i2 = 0; i2 < g.n2; i2 += g.n2 - 1)
        for (int i1 = 0; i1 < g.n1; i1 += g.n1 - 1) {
          const cf corner(g.o2 + i2 * g.d2, g.o1 + i1 * g.d1);
          a = std::max(a, std::abs(corner - b));
        }
      a = 1.f / a;
      const Map sas = shift_scale(g, a, b);
      const Map lg = complex_log(sas.out, eps_log);
      out = lg.out;
      const std::vector<int> fs = forward_gather(sas), fl = log_forward_gather(lg);
      for (size_t i = 0; i < n; ++i) f[i] = fl[i] < 0 ? -1 : fs[(size_t)fl[i]];
      const std::vector<int> is = inverse_gather(sas), il = log_inverse_gather(lg, eps_log);
      if (il.empty()) return fail(WEMFWI_ERR_ARG, "cmap 1: inverse map leaves the grid");
      for (size_t i = 0; i < n; ++i) v[i] = is[i] < 0 ? -1 : il[(size_t)is[i]];
      for (size_t i = 0; i < n; ++i) {
        j1[i] = sas.jac[i];
        j2[i] = lg.jac[i];
      }
      // source: a line along the first row of the log grid (ComplexLog.cpp:167-190)
      nsrc[s] = nx;
      for (int i = 0; i < nx; ++i) {
        sz[i] = 0;
        sx[i] = i;
      }
      // receivers: each through both maps, depth clipped at 0 (ComplexLog.cpp:132-165)
      for (int i = 0; i < nr; ++i) {
        int z1, x1, z2, x2;
        map_point(sas, rec_iz[i], rec_ix[i], z1, x1);
        if (z1 < 0 || z1 >= nz || x1 < 0 || x1 >= nx)
          return fail(WEMFWI_ERR_ARG, "cmap 1: receiver maps outside the grid");
        map_point(lg, z1, x1, z2, x2);
        z2 = std::max(0, z2);
        if (z2 >= nz || x2 < 0 || x2 >= nx) return fail(WEMFWI_ERR_ARG, "cmap 1: receiver maps outside the grid");
        rz[i] = z2;
        rx[i] = x2;
      }
    } else {
      // PropParam.h:111-127: the rotation alone is used
      const Map rt = rotation(g, angles[s]);
      out = rt.out;
      const std::vector<int> fr = forward_gather(rt), ir = inverse_gather(rt);
      for (size_t i = 0; i < n; ++i) {
        f[i] = fr[i];
        v[i] = ir[i];
        j1[i] = 1.f;
        j2[i] = 1.f;
      }
      nsrc[s] = 1;
      map_point(rt, src_iz[s], src_ix[s], sz[0], sx[0]);
      if (sz[0] < 0 || sz[0] >= nz || sx[0] < 0 || sx[0] >= nx)
        return fail(WEMFWI_ERR_ARG, "cmap 2: source maps outside the grid");
      for (int i = 0; i < nr; ++i) {
        map_point(rt, rec_iz[i], rec_ix[i], rz[i], rx[i]);
        if (rz[i] < 0 || rz[i] >= nz || rx[i] < 0 || rx[i] >= nx)
          return fail(WEMFWI_ERR_ARG, "cmap 2: receiver maps outside the grid");
      }
    }
    ax[0] = out.o1;
    ax[1] = out.d1;
    ax[2] = out.o2;
    ax[3] = out.d2;
  }
  return WEMFWI_OK;
}

}  // extern "C"
