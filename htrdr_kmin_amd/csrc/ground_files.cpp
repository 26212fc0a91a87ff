/* Readers of the ground input files of htrdr-atmosphere: htrdr-obj geometry,
 * htrdr-materials lists and mrumtl BRDF files (include/htgx_files.h). Host
 * code only. */
#include "htgx_files.h"

#include <cerrno>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <new>
#include <string>
#include <vector>

namespace {

thread_local std::string g_error;

int fail(int res, const char* fmt, ...)
{
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_error = buf;
  return res;
}

/* Blank-separated tokens of a line, after removing what follows `comment' */
std::vector<std::string> tokens(const std::string& line, char comment)
{
  std::vector<std::string> out;
  const size_t end = comment ? line.find(comment) : std::string::npos;
  const std::string s = line.substr(0, end);
  size_t i = 0;
  while(i < s.size()) {
    while(i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r' || s[i] == '\n')) ++i;
    size_t j = i;
    while(j < s.size() && !(s[j] == ' ' || s[j] == '\t' || s[j] == '\r' || s[j] == '\n')) ++j;
    if(j > i) out.push_back(s.substr(i, j - i));
    i = j;
  }
  return out;
}

/* The whole token as a real number (rsys cstr_to_double semantics) */
bool to_double(const std::string& t, double* out)
{
  if(t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  const double v = std::strtod(t.c_str(), &end);
  if(*end != '\0' || errno == ERANGE) return false;
  *out = v;
  return true;
}

bool to_long(const std::string& t, long* out)
{
  if(t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(t.c_str(), &end, 10);
  if(*end != '\0' || errno == ERANGE) return false;
  *out = v;
  return true;
}

struct Material {
  std::string name;
  bool has_brdf = false;
  int list = HTGX_BRDF_LIST_WAVELENGTHS;
  std::vector<htgx_brdf> brdfs;
  double temperature = 0.0;
};

/* An mrumtl file: a BRDF per wavelength or per band */
int load_mrumtl(const std::string& path, Material& m)
{
  std::ifstream in(path);
  if(!in) return fail(HTGX_IO_ERR, "%s: cannot open the mrumtl file", path.c_str());
  std::string line;
  size_t lnum = 0;
  long count = -1;
  while(std::getline(in, line)) {
    ++lnum;
    const std::vector<std::string> t = tokens(line, '#');
    if(t.empty()) continue;
    if(count < 0) {
      if(t.size() != 2 || (t[0] != "wavelengths" && t[0] != "bands") || !to_long(t[1], &count) || count <= 0)
        return fail(HTGX_BAD_ARG, "%s:%zu: expected `wavelengths <count>' or `bands <count>'",
                    path.c_str(), lnum);
      m.list = t[0] == "bands" ? HTGX_BRDF_LIST_BANDS : HTGX_BRDF_LIST_WAVELENGTHS;
      continue;
    }
    const bool bands = m.list == HTGX_BRDF_LIST_BANDS;
    const size_t nw = bands ? 2 : 1;
    if((long)m.brdfs.size() == count)
      return fail(HTGX_BAD_ARG, "%s:%zu: more entries than the %ld announced", path.c_str(), lnum, count);
    if(t.size() != nw + 2)
      return fail(HTGX_BAD_ARG, "%s:%zu: expected %s <lambertian|specular> <reflectivity>", path.c_str(),
                  lnum, bands ? "<lo> <hi>" : "<wavelength>");
    htgx_brdf e;
    std::memset(&e, 0, sizeof(e));
    if(!to_double(t[0], &e.wlen[0]) || !(e.wlen[0] >= 0))
      return fail(HTGX_BAD_ARG, "%s:%zu: invalid wavelength `%s'", path.c_str(), lnum, t[0].c_str());
    e.wlen[1] = e.wlen[0];
    if(bands && (!to_double(t[1], &e.wlen[1]) || !(e.wlen[1] >= e.wlen[0])))
      return fail(HTGX_BAD_ARG, "%s:%zu: invalid band `%s %s'", path.c_str(), lnum, t[0].c_str(),
                  t[1].c_str());
    const std::string& kind = t[nw];
    if(kind == "lambertian") e.type = HTGX_BRDF_LAMBERTIAN;
    else if(kind == "specular") e.type = HTGX_BRDF_SPECULAR;
    else return fail(HTGX_BAD_ARG, "%s:%zu: unknown BRDF `%s'", path.c_str(), lnum, kind.c_str());
    if(!to_double(t[nw + 1], &e.reflectivity) || !(e.reflectivity >= 0 && e.reflectivity <= 1))
      return fail(HTGX_BAD_ARG, "%s:%zu: invalid reflectivity `%s'", path.c_str(), lnum,
                  t[nw + 1].c_str());
    if(!m.brdfs.empty()) {
      const htgx_brdf& p = m.brdfs.back();
      if(bands ? !(e.wlen[0] >= p.wlen[1]) : !(e.wlen[0] > p.wlen[0]))
        return fail(HTGX_BAD_ARG, "%s:%zu: %s not in increasing order", path.c_str(), lnum,
                    bands ? "bands overlapping or" : "wavelengths");
    }
    m.brdfs.push_back(e);
  }
  if(count < 0) return fail(HTGX_BAD_ARG, "%s: no `wavelengths' or `bands' list", path.c_str());
  if((long)m.brdfs.size() != count)
    return fail(HTGX_BAD_ARG, "%s: %zu entries instead of the %ld announced", path.c_str(),
                m.brdfs.size(), count);
  m.has_brdf = true;
  return HTGX_OK;
}

/* htrdr-materials: `name mrumtl|none [temperature]' per line
 * (src/core/htrdr_materials.c:72-235) */
int load_materials(const char* path, std::vector<Material>& mtls, std::map<std::string, uint32_t>& index)
{
  std::ifstream in(path);
  if(!in) return fail(HTGX_IO_ERR, "%s: cannot open the materials file", path);
  std::string line;
  size_t lnum = 0;
  while(std::getline(in, line)) {
    ++lnum;
    const std::vector<std::string> t = tokens(line, '#');
    if(t.empty()) continue;
    if(t.size() < 2)
      return fail(HTGX_BAD_ARG, "%s:%zu: missing the MruMtl file for the material `%s'", path, lnum,
                  t[0].c_str());
    Material m;
    m.name = t[0];
    if(t[1] != "none") {
      int res = load_mrumtl(t[1], m);
      if(res != HTGX_OK) return res;
      if(t.size() < 3)
        return fail(HTGX_BAD_ARG, "%s:%zu: missing temperature for the material `%s'", path, lnum,
                    t[0].c_str());
    }
    if(t.size() >= 3 && !to_double(t[2], &m.temperature))
      return fail(HTGX_BAD_ARG, "%s:%zu: error parsing the temperature `%s' for the material `%s'",
                  path, lnum, t[2].c_str(), t[0].c_str());
    /* a name listed again replaces its definition */
    auto it = index.find(m.name);
    if(it != index.end()) {
      mtls[it->second] = m;
    } else {
      index[m.name] = (uint32_t)mtls.size();
      mtls.push_back(m);
    }
  }
  return HTGX_OK;
}

/* Index of an obj face vertex `v', `v/vt', `v//vn' or `v/vt/vn' (1-based,
 * negative: relative to the vertices read so far) */
bool face_vertex(const std::string& tok, size_t nverts, uint32_t* out)
{
  const std::string v = tok.substr(0, tok.find('/'));
  long i;
  if(!to_long(v, &i) || i == 0) return false;
  const long k = i > 0 ? i - 1 : (long)nverts + i;
  if(k < 0 || k >= (long)nverts) return false;
  *out = (uint32_t)k;
  return true;
}

/* Materials of a usemtl group name `front[:thin]:back' split as strtok_r does
 * (src/core/htrdr_geometry.c:177-270: empty fields are skipped) */
std::vector<std::string> split_colons(const std::string& s)
{
  std::vector<std::string> out;
  size_t i = 0;
  while(i <= s.size()) {
    const size_t j = s.find(':', i);
    const std::string f = s.substr(i, j == std::string::npos ? std::string::npos : j - i);
    if(!f.empty()) out.push_back(f);
    if(j == std::string::npos) break;
    i = j + 1;
  }
  return out;
}

} /* namespace */

struct htgx_ground_files {
  std::vector<Material> mtls;
  std::vector<std::string> ifc_names;
  std::vector<float> verts;
  std::vector<uint32_t> tris;
  std::vector<uint32_t> tri_ifc;
  std::vector<uint32_t> brdf_first;
  std::vector<int32_t> brdf_list;
  std::vector<htgx_brdf> brdfs;
  std::vector<double> temperature;
  std::vector<int32_t> interfaces;
  htgx_ground_desc desc;
};

extern "C" int
htgx_ground_files_load(const char* obj_path, const char* materials_path, const char* sky_mtl,
                       htgx_ground_files** out)
{
  g_error.clear();
  if(!obj_path || !materials_path || !out) return fail(HTGX_BAD_ARG, "missing argument");
  htgx_ground_files* f = new(std::nothrow) htgx_ground_files();
  if(!f) return HTGX_MEM_ERR;
  auto done = [&](int res) {
    if(res != HTGX_OK) { delete f; f = nullptr; }
    *out = f;
    return res;
  };

  std::map<std::string, uint32_t> mindex;
  int res = load_materials(materials_path, f->mtls, mindex);
  if(res != HTGX_OK) return done(res);

  std::ifstream in(obj_path);
  if(!in) return done(fail(HTGX_IO_ERR, "%s: cannot open the obj file", obj_path));
  std::map<std::string, uint32_t> iindex;
  std::string line;
  size_t lnum = 0, nusemtl = 0;
  int64_t cur = -1; /* interface of the current usemtl group; -1 before the first */
  size_t nverts = 0;
  while(std::getline(in, line)) {
    ++lnum;
    const std::vector<std::string> t = tokens(line, '#');
    if(t.empty()) continue;
    if(t[0] == "v") {
      double x[3];
      /* x y z, then an optional w or vertex colour, ignored */
      if(t.size() < 4 || !to_double(t[1], &x[0]) || !to_double(t[2], &x[1]) ||
         !to_double(t[3], &x[2]))
        return done(fail(HTGX_BAD_ARG, "%s:%zu: invalid vertex", obj_path, lnum));
      for(int a = 0; a < 3; ++a) f->verts.push_back((float)x[a]);
      ++nverts;
    } else if(t[0] == "f") {
      if(t.size() != 4)
        return done(fail(HTGX_BAD_ARG,
          "%s:%zu: the obj has non-triangulated polygons while only triangular meshes are supported",
          obj_path, lnum));
      uint32_t v[3];
      for(int k = 0; k < 3; ++k)
        if(!face_vertex(t[1 + k], nverts, &v[k]))
          return done(fail(HTGX_BAD_ARG, "%s:%zu: invalid face vertex `%s'", obj_path, lnum,
                           t[1 + k].c_str()));
      if(cur < 0) continue; /* in no material group: not part of the ground */
      for(int k = 0; k < 3; ++k) f->tris.push_back(v[k]);
      f->tri_ifc.push_back((uint32_t)cur);
    } else if(t[0] == "usemtl") {
      if(t.size() != 2)
        return done(fail(HTGX_BAD_ARG, "%s:%zu: invalid usemtl directive", obj_path, lnum));
      ++nusemtl;
      auto it = iindex.find(t[1]);
      if(it != iindex.end()) { cur = it->second; continue; }
      const std::vector<std::string> names = split_colons(t[1]);
      if(names.size() < 2)
        return done(fail(HTGX_BAD_ARG, "%s:%zu: the back material name is missing `%s'", obj_path,
                         lnum, t[1].c_str()));
      /* front[:thin]:back */
      const std::string& front = names[0];
      const std::string* thin = names.size() >= 3 ? &names[1] : nullptr;
      const std::string& back = names.size() >= 3 ? names[2] : names[1];
      int32_t ids[3] = {-1, -1, -1};
      const std::string* want[3] = {&front, &back, thin};
      const char* what[3] = {"front", "back", "interface"};
      for(int k = 0; k < 3; ++k) {
        if(!want[k]) continue;
        auto m = mindex.find(*want[k]);
        if(m == mindex.end())
          return done(fail(HTGX_BAD_ARG, "%s:%zu: invalid interface `%s': the %s material `%s' is unknown",
                           obj_path, lnum, t[1].c_str(), what[k], want[k]->c_str()));
        ids[k] = (int32_t)m->second;
      }
      const bool reflects = f->mtls[ids[0]].has_brdf || f->mtls[ids[1]].has_brdf ||
                            (ids[2] >= 0 && f->mtls[ids[2]].has_brdf);
      if(!reflects)
        return done(fail(HTGX_BAD_ARG, "%s:%zu: interface `%s': none of its materials has a BRDF",
                         obj_path, lnum, t[1].c_str()));
      cur = (int64_t)f->ifc_names.size();
      iindex[t[1]] = (uint32_t)cur;
      f->ifc_names.push_back(t[1]);
      for(int k = 0; k < 3; ++k) f->interfaces.push_back(ids[k]);
    }
    /* vt, vn, vp, g, o, s, mtllib, l, p: no bearing on the ground */
  }
  if(!nusemtl) return done(fail(HTGX_BAD_ARG, "%s: the obj has no material", obj_path));
  if(f->tris.empty()) return done(fail(HTGX_BAD_ARG, "%s: no triangle in a material group", obj_path));

  /* the description */
  f->brdf_first.push_back(0);
  for(const Material& m : f->mtls) {
    f->brdf_list.push_back(m.list);
    f->brdfs.insert(f->brdfs.end(), m.brdfs.begin(), m.brdfs.end());
    f->brdf_first.push_back((uint32_t)f->brdfs.size());
    f->temperature.push_back(m.temperature);
  }
  const std::string sky = sky_mtl ? sky_mtl : "air";
  auto s = mindex.find(sky);
  htgx_ground_desc& d = f->desc;
  std::memset(&d, 0, sizeof(d));
  d.type = HTGX_GROUND_MESH;
  d.nverts = (uint32_t)nverts;
  d.ntris = (uint32_t)f->tri_ifc.size();
  d.verts = f->verts.data();
  d.tris = f->tris.data();
  d.tri_mtl = f->tri_ifc.data();
  d.nmtls = (uint32_t)f->mtls.size();
  d.mtl_temperature = f->temperature.data();
  d.mtl_brdf_first = f->brdf_first.data();
  d.mtl_brdf_list = f->brdf_list.data();
  d.brdfs = f->brdfs.empty() ? nullptr : f->brdfs.data();
  d.ninterfaces = (uint32_t)f->ifc_names.size();
  d.interfaces = f->interfaces.data();
  d.sky_mtl = s == mindex.end() ? -1 : (int32_t)s->second;
  return done(HTGX_OK);
}

extern "C" const htgx_ground_desc*
htgx_ground_files_get_desc(const htgx_ground_files* f)
{
  return f ? &f->desc : nullptr;
}

extern "C" const char*
htgx_ground_files_material_name(const htgx_ground_files* f, uint32_t m)
{
  return f && m < f->mtls.size() ? f->mtls[m].name.c_str() : nullptr;
}

extern "C" const char*
htgx_ground_files_interface_name(const htgx_ground_files* f, uint32_t i)
{
  return f && i < f->ifc_names.size() ? f->ifc_names[i].c_str() : nullptr;
}

extern "C" int
htgx_ground_files_destroy(htgx_ground_files* f)
{
  delete f;
  return HTGX_OK;
}

extern "C" const char*
htgx_files_last_error(void)
{
  return g_error.c_str();
}
