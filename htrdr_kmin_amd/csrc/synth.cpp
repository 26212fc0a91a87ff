/*
 * Seeded synthetic atmospheres (the inputs htsky would read from htgop, htcp
 * and htmie files, src/atmosphere/htrdr_atmosphere.c:366-382).
 */
#include "htgx_synth.h"
#include "htgx_rng.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

struct htgx_synth {
  htgx_scene_desc scene;
  std::vector<double> band_lo, band_hi, quad_weights;
  std::vector<uint32_t> nquads;
  std::vector<double> ka, ks, temperature;
  std::vector<float> lwc;
  std::vector<double> cext, ssa, hg_g;
  std::vector<double> mtl_albedo, mtl_temperature;
  std::vector<float> verts;
  std::vector<uint32_t> tris, tri_mtl;
};

namespace {

const double kPi = 3.14159265358979323846;

/* Fill `field' (nx*ny*nz, X fastest) with N(0,1) numbers. Each Z slice owns
 * the Threefry stream (seed, field_id, z) and draws Box-Muller pairs. */
void gaussian_field(uint64_t seed, uint64_t field_id, const uint32_t def[3],
                    std::vector<float>& field)
{
  const size_t nxy = (size_t)def[0] * def[1];
  field.resize(nxy * def[2]);
  #pragma omp parallel for schedule(static)
  for(int64_t z = 0; z < (int64_t)def[2]; ++z) {
    htgx::Stream rng;
    rng.init(seed, field_id, (uint64_t)z);
    float* out = field.data() + (size_t)z * nxy;
    for(size_t i = 0; i < nxy; i += 2) {
      const double u0 = rng.canonical();
      const double u1 = rng.canonical();
      const double r = std::sqrt(-2.0 * std::log(1.0 - u0));
      out[i] = (float)(r * std::cos(2.0 * kPi * u1));
      if(i + 1 < nxy) out[i + 1] = (float)(r * std::sin(2.0 * kPi * u1));
    }
  }
}

/* Periodic separable Gaussian smoothing, kernel truncated at 4 sigma */
void smooth(std::vector<float>& f, const uint32_t def[3], double sigma)
{
  const int R = (int)std::ceil(4.0 * sigma);
  std::vector<double> w(2 * R + 1);
  double sum = 0;
  for(int k = -R; k <= R; ++k) {
    w[k + R] = std::exp(-0.5 * k * k / (sigma * sigma));
    sum += w[k + R];
  }
  for(double& v : w) v /= sum;

  const size_t nx = def[0], ny = def[1], nz = def[2];
  std::vector<float> tmp(f.size());
  const size_t stride[3] = {1, nx, nx * ny};
  const size_t n[3] = {nx, ny, nz};
  for(int axis = 0; axis < 3; ++axis) {
    const size_t na = n[axis];
    const size_t st = stride[axis];
    #pragma omp parallel for schedule(static)
    for(int64_t idx = 0; idx < (int64_t)f.size(); ++idx) {
      const size_t ia = ((size_t)idx / st) % na;
      const size_t base = (size_t)idx - ia * st;
      double acc = 0;
      for(int k = -R; k <= R; ++k) {
        int64_t j = (int64_t)ia + k;
        j %= (int64_t)na;
        if(j < 0) j += (int64_t)na;
        acc += w[k + R] * (double)f[base + (size_t)j * st];
      }
      tmp[(size_t)idx] = (float)acc;
    }
    f.swap(tmp);
  }
}

/* Value of the periodic lattice noise of octave `oct' (n x n points) at
 * (x, y) in [0, 1)^2: smoothstep interpolation of per-point values uniform in
 * [-1, 1], drawn from the Threefry stream (seed, 0x6866 "hf", oct) */
double value_noise(const std::vector<double>& lat, uint32_t n, double x, double y)
{
  const double fx = x * n, fy = y * n;
  const uint32_t i0 = (uint32_t)fx % n, j0 = (uint32_t)fy % n;
  const uint32_t i1 = (i0 + 1) % n, j1 = (j0 + 1) % n;
  double tx = fx - std::floor(fx), ty = fy - std::floor(fy);
  tx = tx * tx * (3 - 2 * tx);
  ty = ty * ty * (3 - 2 * ty);
  const double a = lat[(size_t)j0 * n + i0], b = lat[(size_t)j0 * n + i1];
  const double c = lat[(size_t)j1 * n + i0], d = lat[(size_t)j1 * n + i1];
  return (a * (1 - tx) + b * tx) * (1 - ty) + (c * (1 - tx) + d * tx) * ty;
}

/* fBm heightfield mesh over [x0, x1] x [y0, y1] */
void heightfield(const htgx_synth_params* p, double x0, double x1, double y0, double y1,
                 htgx_synth* s)
{
  const uint32_t R = p->hf_res;
  std::vector<double> h((size_t)R * R, 0.0);
  for(uint32_t o = 0; o < p->hf_octaves; ++o) {
    const uint32_t n = 2u << o; /* lattice points per side */
    std::vector<double> lat((size_t)n * n);
    htgx::Stream rng;
    rng.init(p->seed, 0x6866ULL, o);
    for(double& v : lat) v = 2.0 * rng.canonical() - 1.0;
    const double amp = std::pow(2.0, -p->hf_hurst * (double)o);
    #pragma omp parallel for schedule(static)
    for(int64_t j = 0; j < (int64_t)R; ++j)
      for(uint32_t i = 0; i < R; ++i)
        h[(size_t)j * R + i] += amp * value_noise(lat, n, (double)i / R, (double)j / R);
  }
  const auto mm = std::minmax_element(h.begin(), h.end());
  const double lo = *mm.first, span = *mm.second - *mm.first;
  s->verts.resize(3 * (size_t)R * R);
  for(uint32_t j = 0; j < R; ++j)
    for(uint32_t i = 0; i < R; ++i) {
      float* v = &s->verts[3 * ((size_t)j * R + i)];
      v[0] = (float)(x0 + (x1 - x0) * i / (R - 1));
      v[1] = (float)(y0 + (y1 - y0) * j / (R - 1));
      v[2] = (float)(span > 0 ? (h[(size_t)j * R + i] - lo) / span * p->hf_relief : 0.0);
    }
  for(uint32_t j = 0; j + 1 < R; ++j)
    for(uint32_t i = 0; i + 1 < R; ++i) {
      const uint32_t v00 = j * R + i, v10 = v00 + 1, v01 = v00 + R, v11 = v01 + 1;
      const uint32_t quad[2][3] = {{v00, v10, v11}, {v00, v11, v01}};
      for(const auto& t : quad) {
        const double zc = (s->verts[3 * t[0] + 2] + s->verts[3 * t[1] + 2] + s->verts[3 * t[2] + 2]) / 3.0;
        const uint32_t band = p->hf_relief > 0
          ? std::min(3u, (uint32_t)(4.0 * zc / p->hf_relief)) : 0u;
        s->tris.insert(s->tris.end(), {t[0], t[1], t[2]});
        s->tri_mtl.push_back(band);
      }
    }
}

} /* namespace */

extern "C" int
htgx_synth_params_preset(int config, htgx_synth_params* p)
{
  if(!p) return HTGX_BAD_ARG;
  std::memset(p, 0, sizeof(*p));
  p->spectral_type = HTGX_SPECTRAL_SW;
  p->lwc_mu = -1.6;
  p->lwc_sigma = 0.8;
  p->cover = 0.5;
  p->smooth_sigma = 4.0;
  p->cext = 0.15;      /* 150 m^2/kg */
  p->ssa = 0.999;
  p->hg_g = 0.85;
  p->nlayers = 50;
  p->layer_thickness = 400.0;
  p->ka0 = 1e-5;
  p->ks0 = 1.2e-5;
  p->scale_height = 8000.0;
  p->T_ground = 288.0;
  p->lapse_rate = 6.5e-3;
  p->nbands = 1;
  p->nquads = 1;
  p->wlen_lo = 500.0;
  p->wlen_hi = 600.0;
  p->lw_ka_min = 1e-7;
  p->lw_ka_max = 1e-2;
  p->ground_type = HTGX_GROUND_PLANE;
  p->ground_albedo = 0.3;
  p->ground_temperature = 288.0;
  p->sun_azimuth = 0.0;
  p->sun_elevation = 60.0; /* zenith 30 deg */
  p->optical_thickness = 1.0;
  p->leaf_block = 2;
  p->seed = 0;
  switch(config) {
    case 0:
      p->cloud_def[0] = 64; p->cloud_def[1] = 64; p->cloud_def[2] = 32;
      p->cloud_upper[0] = 6400; p->cloud_upper[1] = 6400;
      break;
    case 1:
      p->cloud_def[0] = 256; p->cloud_def[1] = 256; p->cloud_def[2] = 64;
      p->cloud_upper[0] = 25600; p->cloud_upper[1] = 25600;
      break;
    case 2:
      p->cloud_def[0] = 256; p->cloud_def[1] = 256; p->cloud_def[2] = 64;
      p->cloud_upper[0] = 25600; p->cloud_upper[1] = 25600;
      p->spectral_type = HTGX_SPECTRAL_LW;
      p->nbands = 8;
      p->nquads = 16;
      p->wlen_lo = 4000.0;
      p->wlen_hi = 20000.0;
      break;
    case 4:
      p->cloud_def[0] = 512; p->cloud_def[1] = 512; p->cloud_def[2] = 128;
      p->cloud_upper[0] = 51200; p->cloud_upper[1] = 51200;
      p->ground_type = HTGX_GROUND_MESH;
      p->hf_res = 1024;
      p->hf_octaves = 8;
      p->hf_relief = 800.0;
      p->hf_hurst = 0.7;
      break;
    default: return HTGX_BAD_ARG;
  }
  p->cloud_lower[2] = 1000.0;
  p->cloud_upper[2] = config == 4 ? 7400.0 : 4200.0;
  return HTGX_OK;
}

extern "C" int
htgx_synth_map_preset(int config, htgx_map_desc* d)
{
  if(!d) return HTGX_BAD_ARG;
  std::memset(d, 0, sizeof(*d));
  const double tilt = 60.0 * kPi / 180.0; /* camera tilted 60 deg down */
  d->up[2] = 1.0;
  d->fov_y = 60.0;
  d->seed = 0;
  d->solver = HTGX_SOLVER_PLAIN_FP64;
  switch(config) {
    case 0:
    case 1:
      d->width = d->height = config == 0 ? 64 : 1024;
      d->spp = config == 0 ? 32 : 256;
      d->sensor = HTGX_SENSOR_CAMERA;
      d->pos[0] = config == 0 ? 3200.0 : 12800.0;
      d->pos[1] = -2000.0;
      d->pos[2] = 5000.0;
      d->tgt[0] = d->pos[0];
      d->tgt[1] = d->pos[1] + std::cos(tilt);
      d->tgt[2] = d->pos[2] - std::sin(tilt);
      d->spectral_type = HTGX_SPECTRAL_SW;
      d->wlen_range[0] = d->wlen_range[1] = 550.0;
      d->ref_temperature = 5778.0;
      break;
    case 2:
      d->width = d->height = 512;
      d->spp = 256;
      d->sensor = HTGX_SENSOR_RECTANGLE;
      d->pos[0] = 12800.0; d->pos[1] = 12800.0; d->pos[2] = 2000.0;
      d->tgt[0] = 12800.0; d->tgt[1] = 12800.0; d->tgt[2] = 2001.0;
      d->up[0] = 0; d->up[1] = 1; d->up[2] = 0;
      d->rect_size[0] = d->rect_size[1] = 25600.0;
      d->spectral_type = HTGX_SPECTRAL_LW;
      d->wlen_range[0] = 4000.0;
      d->wlen_range[1] = 20000.0;
      d->ref_temperature = 290.0;
      break;
    case 4:
      /* 512^2 SW flux map at 2 km over the whole 51.2 km domain */
      d->width = d->height = 512;
      d->spp = 512;
      d->sensor = HTGX_SENSOR_RECTANGLE;
      d->pos[0] = 25600.0; d->pos[1] = 25600.0; d->pos[2] = 2000.0;
      d->tgt[0] = 25600.0; d->tgt[1] = 25600.0; d->tgt[2] = 2001.0;
      d->up[0] = 0; d->up[1] = 1; d->up[2] = 0;
      d->rect_size[0] = d->rect_size[1] = 51200.0;
      d->spectral_type = HTGX_SPECTRAL_SW;
      d->wlen_range[0] = d->wlen_range[1] = 550.0;
      d->ref_temperature = 5778.0;
      break;
    default: return HTGX_BAD_ARG;
  }
  return HTGX_OK;
}

extern "C" int
htgx_synth_create(const htgx_synth_params* p, htgx_synth** out)
{
  if(!p || !out) return HTGX_BAD_ARG;
  if(!p->nlayers || !(p->layer_thickness > 0) || !p->nbands || !p->nquads
  || !(p->wlen_hi >= p->wlen_lo) || !p->leaf_block) return HTGX_BAD_ARG;
  for(int a = 0; a < 3; ++a) {
    if(p->cloud_def[a] % p->leaf_block) return HTGX_BAD_ARG;
    if(p->cloud_def[a] && !(p->cloud_upper[a] > p->cloud_lower[a])) return HTGX_BAD_ARG;
  }
  htgx_synth* s = new(std::nothrow) htgx_synth();
  if(!s) return HTGX_MEM_ERR;
  std::memset(&s->scene, 0, sizeof(s->scene));
  const uint32_t nb = p->nbands, nq = p->nquads, nl = p->nlayers;

  /* Spectral bands of equal width over [wlen_lo, wlen_hi] */
  const double bw = (p->wlen_hi - p->wlen_lo) / nb;
  for(uint32_t b = 0; b < nb; ++b) {
    s->band_lo.push_back(p->wlen_lo + b * bw);
    s->band_hi.push_back(b + 1 == nb ? p->wlen_hi : p->wlen_lo + (b + 1) * bw);
    s->nquads.push_back(nq);
  }

  /* Gas layers */
  s->ka.resize((size_t)nb * nq * nl);
  s->ks.resize((size_t)nb * nl);
  s->temperature.resize(nl);
  htgx::Stream rng;
  rng.init(p->seed, 0x6761735f6b61ULL /* spectral stream */, 0);
  for(uint32_t b = 0; b < nb; ++b) {
    std::vector<double> w(nq);
    double wsum = 0;
    for(uint32_t q = 0; q < nq; ++q) {
      double ka0 = p->ka0;
      if(p->spectral_type == HTGX_SPECTRAL_LW) {
        const double lmin = std::log(p->lw_ka_min), lmax = std::log(p->lw_ka_max);
        ka0 = std::exp(lmin + (lmax - lmin) * rng.canonical());
        w[q] = -std::log(1.0 - rng.canonical()); /* Dirichlet(1) */
      } else if(nq > 1) {
        /* several shortwave k-points: ka0 log-uniform within a decade */
        ka0 = p->ka0 * std::pow(10.0, 2.0 * rng.canonical() - 1.0);
        w[q] = -std::log(1.0 - rng.canonical());
      } else {
        w[q] = 1.0;
      }
      wsum += w[q];
      for(uint32_t l = 0; l < nl; ++l) {
        const double z = (l + 0.5) * p->layer_thickness;
        s->ka[((size_t)b * nq + q) * nl + l] = ka0 * std::exp(-z / p->scale_height);
      }
    }
    for(uint32_t q = 0; q < nq; ++q) s->quad_weights.push_back(w[q] / wsum);
    for(uint32_t l = 0; l < nl; ++l) {
      const double z = (l + 0.5) * p->layer_thickness;
      s->ks[(size_t)b * nl + l] = p->spectral_type == HTGX_SPECTRAL_LW
        ? 0.0 : p->ks0 * std::exp(-z / p->scale_height);
    }
  }
  for(uint32_t l = 0; l < nl; ++l) {
    const double z = (l + 0.5) * p->layer_thickness;
    s->temperature[l] = p->T_ground - p->lapse_rate * z;
  }

  /* Clouds */
  const bool clouds = p->cloud_def[0] && p->cloud_def[1] && p->cloud_def[2];
  if(clouds) {
    std::vector<float> lwc_n, mask_n;
    gaussian_field(p->seed, 1, p->cloud_def, lwc_n);
    gaussian_field(p->seed, 2, p->cloud_def, mask_n);
    smooth(mask_n, p->cloud_def, p->smooth_sigma);
    std::vector<float> sorted(mask_n);
    const size_t n = sorted.size();
    size_t k = (size_t)((1.0 - p->cover) * (double)n);
    if(k >= n) k = n - 1;
    std::nth_element(sorted.begin(), sorted.begin() + (ptrdiff_t)k, sorted.end());
    const float threshold = sorted[k];
    s->lwc.resize(n);
    for(size_t i = 0; i < n; ++i) {
      s->lwc[i] = mask_n[i] >= threshold && p->cover > 0
        ? (float)std::exp(p->lwc_mu + p->lwc_sigma * (double)lwc_n[i]) : 0.f;
    }
    s->cext.assign(nb, p->cext);
    s->ssa.assign(nb, p->ssa);
    s->hg_g.assign(nb, p->hg_g);
  }
  if(p->ground_type == HTGX_GROUND_MESH) {
    if(p->hf_res < 2 || !p->hf_octaves || p->hf_octaves > 20) { delete s; return HTGX_BAD_ARG; }
    /* over the configured cloud extent (also when the clouds are off) */
    const bool ext = p->cloud_upper[0] > p->cloud_lower[0] && p->cloud_upper[1] > p->cloud_lower[1];
    const double x0 = ext ? p->cloud_lower[0] : 0.0, x1 = ext ? p->cloud_upper[0] : 1.0;
    const double y0 = ext ? p->cloud_lower[1] : 0.0, y1 = ext ? p->cloud_upper[1] : 1.0;
    heightfield(p, x0, x1, y0, y1, s);
    s->mtl_albedo = {0.1, 0.25, 0.4, 0.6};
    s->mtl_temperature.assign(4, p->ground_temperature);
  } else {
    s->mtl_albedo.push_back(p->ground_albedo);
    s->mtl_temperature.push_back(p->ground_temperature);
  }

  htgx_scene_desc& sc = s->scene;
  sc.spectral.nbands = nb;
  sc.spectral.band_lo = s->band_lo.data();
  sc.spectral.band_hi = s->band_hi.data();
  sc.spectral.nquads = s->nquads.data();
  sc.spectral.quad_weights = s->quad_weights.data();
  sc.gas.nlayers = nl;
  sc.gas.z_bottom = 0.0;
  sc.gas.layer_thickness = p->layer_thickness;
  sc.gas.ka = s->ka.data();
  sc.gas.ks = s->ks.data();
  sc.gas.temperature = s->temperature.data();
  if(clouds) {
    for(int a = 0; a < 3; ++a) {
      sc.clouds.def[a] = p->cloud_def[a];
      sc.clouds.lower[a] = p->cloud_lower[a];
      sc.clouds.upper[a] = p->cloud_upper[a];
    }
    sc.clouds.lwc = s->lwc.data();
    sc.clouds.cext = s->cext.data();
    sc.clouds.ssa = s->ssa.data();
    sc.clouds.hg_g = s->hg_g.data();
  }
  sc.ground.type = p->ground_type;
  sc.ground.plane_z = 0.0;
  sc.ground.plane_mtl = 0;
  sc.ground.nmtls = (uint32_t)s->mtl_albedo.size();
  if(p->ground_type == HTGX_GROUND_MESH) {
    sc.ground.nverts = (uint32_t)(s->verts.size() / 3);
    sc.ground.ntris = (uint32_t)(s->tris.size() / 3);
    sc.ground.verts = s->verts.data();
    sc.ground.tris = s->tris.data();
    sc.ground.tri_mtl = s->tri_mtl.data();
  }
  sc.ground.mtl_albedo = s->mtl_albedo.data();
  sc.ground.mtl_temperature = s->mtl_temperature.data();
  sc.sun_azimuth = p->sun_azimuth;
  sc.sun_elevation = p->sun_elevation;
  sc.optical_thickness = p->optical_thickness;
  sc.leaf_block = p->leaf_block;
  *out = s;
  return HTGX_OK;
}

extern "C" const htgx_scene_desc*
htgx_synth_get_scene(const htgx_synth* s)
{
  return s ? &s->scene : nullptr;
}

extern "C" int
htgx_synth_destroy(htgx_synth* s)
{
  delete s;
  return HTGX_OK;
}
