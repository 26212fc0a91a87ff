/*
 * htrdr-atmosphere over libhtgx.so: the command line of the reference
 * (src/atmosphere/htrdr_atmosphere_args.c:124-255, same getopt string and
 * sub-option syntax), driving htgx_draw_map in place of htrdr_draw_map
 * (src/atmosphere/htrdr_atmosphere_draw_map.c:593-641) and writing the same
 * htrdr-image (doc/htrdr-image.5) as dump_buffer (:477-587).
 *
 * The ground of -g is an htrdr-obj file with the htrdr-materials list of -M
 * and the sky material of -n, read by htgx_ground_files_load
 * (include/htgx_files.h). The atmosphere inputs, the htgop (-a), htcp (-c)
 * and htmie (-m) files, are read by htsky in the reference; no reader for
 * these formats exists in this library, so those options take a synthetic
 * source instead of a path:
 *   -a synth:<config>[,<seed>]   gas layers and spectral data of a seeded
 *                                benchmark configuration (0, 1, 2, 4)
 *   -c synth                     its cloud field (the Mie data come with it,
 *                                so -m is not required)
 * and -g also takes
 *   plane | synth                an infinite plane (the quad repeated with -R)
 *                                or the configuration's fBm heightfield.
 * -d writes the octrees (htgx_dump_octrees_vtk) instead of rendering.
 * -O reads the octree from the cache file, or writes it there when the file
 * does not exist (htgx_ctx_create_cached); a cache built from other inputs
 * aborts the run.
 * -V x,y,z bounds the definition of the octree's finest level: the leaf
 * block becomes the smallest cube of cloud cells that divides the cloud grid
 * and keeps at most x,y,z blocks along each axis
 * (htgx_leaf_block_for_definition); ignored when a cache is read (-O).
 * Options without a counterpart on the GPU path are rejected with an error
 * (-C focal-length) or accepted and ignored (-t).
 * The solver is the FP32 wavefront one; HTGX_SOLVER=plain selects the FP64
 * plain one.
 */
#define _POSIX_C_SOURCE 200809L
#include "htgx.h"
#include "htgx_files.h"
#include "htgx_synth.h"

#include <errno.h>
#include <fcntl.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

struct args {
  const char* gas;
  const char* clouds;
  const char* mie;
  const char* ground;
  const char* mtl;
  const char* sky_mtl;
  const char* output;
  int force, verbose, repeat_ground, repeat_clouds, dump;
  int output_type; /* 0 image (-C), 1 flux map (-p) */
  double pos[3], tgt[3], up[3], fov; /* camera */
  int ortho;                         /* -P: orthographic camera */
  double lens_radius, focal_dst;     /* -C thin lens */
  double ortho_height;               /* -P image plane height */
  double rpos[3], rtgt[3], rup[3], rsz[2]; /* flux map rectangle */
  unsigned def[2], spp;
  int spectral; /* enum htgx_spectral_type */
  double wlen[2], tref;
  double sun_az, sun_el;
  double thickness;
  unsigned vdef[3];  /* -V: maximum definition of the octree (0: the clouds') */
  const char* cache; /* -O: octree cache file */
};

static void
usage(void)
{
  printf(
"Usage: htrdr-atmosphere [-dfhRrv] [-s spectral_domain] [-i image]\n"
"                        [-C persp_camera | -P ortho_camera | -p flux_map]\n"
"                        [-D azimuth,elevation] [-T threshold] [-V x,y,z]\n"
"                        [-g ground [-M materials]] [-n sky_mtl]\n"
"                        [-c clouds [-m mie]] [-O cache] [-o output] [-t threads]\n"
"                        -a gas\n"
"Sources without file readers on this GPU path: -a synth:<config>[,<seed>],\n"
"-c synth; -g also takes plane|synth. HTGX_SOLVER=plain selects the FP64 solver.\n");
}

static int
parse_doubles(const char* s, double* out, int n)
{
  char* end;
  int i;
  for(i = 0; i < n; ++i) {
    errno = 0;
    out[i] = strtod(s, &end);
    if(end == s || errno) return -1;
    s = end;
    if(i + 1 < n) {
      if(*s != ',') return -1;
      ++s;
    }
  }
  return *s ? -1 : 0;
}

/* -V x,y,z (parse_grid_definition, src/atmosphere/htrdr_atmosphere_args.c:48-78) */
static int
parse_grid_definition(const char* s, unsigned def[3])
{
  char* end;
  int i;
  for(i = 0; i < 3; ++i) {
    unsigned long v;
    errno = 0;
    v = strtoul(s, &end, 10);
    if(end == s || errno || v > 0xFFFFFFFFul || *s == '-') {
      fprintf(stderr, "Invalid grid definition `%s'.\n", s);
      return -1;
    }
    def[i] = (unsigned)v;
    s = end;
    if(i < 2) {
      if(*s != ',') { fprintf(stderr, "Invalid grid definition `%s'.\n", s); return -1; }
      ++s;
    }
  }
  if(*s) { fprintf(stderr, "Invalid grid definition `%s'.\n", s); return -1; }
  if(!def[0] || !def[1] || !def[2]) {
    fprintf(stderr, "Invalid null grid definition {%u, %u, %u}.\n", def[0], def[1], def[2]);
    return -1;
  }
  return 0;
}

/* "key=value:key=value..." with the handler of one parameter */
typedef int (*param_fn)(const char* key, const char* val, struct args* a);

static int
parse_list(const char* str, param_fn fn, struct args* a)
{
  char buf[512];
  char* ctx = NULL;
  char* tok;
  if(strlen(str) >= sizeof(buf)) return -1;
  strcpy(buf, str);
  for(tok = strtok_r(buf, ":", &ctx); tok; tok = strtok_r(NULL, ":", &ctx)) {
    char* eq = strchr(tok, '=');
    if(eq) *eq = '\0';
    if(fn(tok, eq ? eq + 1 : NULL, a)) return -1;
  }
  return 0;
}

/* -C: htrdr_args_camera_perspective_parse (src/core/htrdr_args.c:255-300) */
static int
camera_param(const char* key, const char* val, struct args* a)
{
  if(!val) return -1;
  if(!strcmp(key, "pos")) return parse_doubles(val, a->pos, 3);
  if(!strcmp(key, "tgt")) return parse_doubles(val, a->tgt, 3);
  if(!strcmp(key, "up")) return parse_doubles(val, a->up, 3);
  if(!strcmp(key, "fov")) {
    if(parse_doubles(val, &a->fov, 1)) return -1;
    return a->fov > 0 && a->fov < 180 ? 0 : -1;
  }
  if(!strcmp(key, "lens-radius")) {
    if(parse_doubles(val, &a->lens_radius, 1)) return -1;
    return a->lens_radius >= 0 ? 0 : -1;
  }
  if(!strcmp(key, "focal-dst")) {
    if(parse_doubles(val, &a->focal_dst, 1)) return -1;
    return a->focal_dst > 0 ? 0 : -1;
  }
  if(!strcmp(key, "focal-length")) {
    fprintf(stderr, "focal-length is not supported on this path: give the field of view (fov)\n");
    return -1;
  }
  return -1;
}

/* -P: htrdr_args_camera_orthographic_parse (src/core/htrdr_args.c:303-345) */
static int
ortho_param(const char* key, const char* val, struct args* a)
{
  if(!val) return -1;
  if(!strcmp(key, "pos")) return parse_doubles(val, a->pos, 3);
  if(!strcmp(key, "tgt")) return parse_doubles(val, a->tgt, 3);
  if(!strcmp(key, "up")) return parse_doubles(val, a->up, 3);
  if(!strcmp(key, "height")) {
    if(parse_doubles(val, &a->ortho_height, 1)) return -1;
    return a->ortho_height > 0 ? 0 : -1;
  }
  return -1;
}

/* -p: htrdr_args_rectangle_parse (src/core/htrdr_args.c:370-410) */
static int
rect_param(const char* key, const char* val, struct args* a)
{
  if(!val) return -1;
  if(!strcmp(key, "pos")) return parse_doubles(val, a->rpos, 3);
  if(!strcmp(key, "tgt")) return parse_doubles(val, a->rtgt, 3);
  if(!strcmp(key, "up")) return parse_doubles(val, a->rup, 3);
  if(!strcmp(key, "sz")) return parse_doubles(val, a->rsz, 2);
  return -1;
}

/* -i: htrdr_args_image_parse (src/core/htrdr_args.c:173-230) */
static int
image_param(const char* key, const char* val, struct args* a)
{
  if(!val) return -1;
  if(!strcmp(key, "def")) {
    if(sscanf(val, "%ux%u", &a->def[0], &a->def[1]) != 2) return -1;
    return a->def[0] && a->def[1] && a->def[0] <= 16384 && a->def[1] <= 16384 ? 0 : -1;
  }
  if(!strcmp(key, "spp")) {
    char* end;
    const unsigned long v = strtoul(val, &end, 10);
    if(*end || !v || v > 0xFFFFFFFFul) return -1;
    a->spp = (unsigned)v;
    return 0;
  }
  return -1;
}

/* -s: htrdr_args_spectral_parse (src/core/htrdr_args.c:440-480) */
static int
spectral_param(const char* key, const char* val, struct args* a)
{
  if(!strcmp(key, "cie_xyz")) {
    /* the CIE 1931 colour space, src/core/htrdr_args.c:451-455 */
    a->spectral = HTGX_SPECTRAL_SW_CIE_XYZ;
    a->wlen[0] = 380;
    a->wlen[1] = 780;
    return 0;
  }
  if(!val) return -1;
  if(!strcmp(key, "sw") || !strcmp(key, "lw")) {
    a->spectral = key[0] == 's' ? HTGX_SPECTRAL_SW : HTGX_SPECTRAL_LW;
    if(parse_doubles(val, a->wlen, 2)) return -1;
    return a->wlen[0] <= a->wlen[1] && a->wlen[0] > 0 ? 0 : -1;
  }
  if(!strcmp(key, "Tref")) {
    if(parse_doubles(val, &a->tref, 1)) return -1;
    return a->tref >= 0 ? 0 : -1;
  }
  return -1;
}

static int
parse_args(int argc, char** argv, struct args* a)
{
  int opt;
  memset(a, 0, sizeof(*a));
  /* defaults of the reference (config.mk:36-53) */
  a->tgt[1] = 1; a->up[2] = 1; a->fov = 70;
  a->focal_dst = 1; a->ortho_height = 1;
  a->rtgt[2] = 1; a->rup[1] = 1; a->rsz[0] = a->rsz[1] = 1;
  a->def[0] = 320; a->def[1] = 240; a->spp = 1;
  a->spectral = HTGX_SPECTRAL_SW_CIE_XYZ;
  a->wlen[0] = 380; a->wlen[1] = 780;
  a->tref = -1;
  a->sun_el = 90;
  a->thickness = 1;
  while((opt = getopt(argc, argv, "a:C:c:D:dfg:hi:M:m:n:O:o:P:p:Rrs:T:t:V:v")) != -1) {
    int bad = 0;
    switch(opt) {
      case 'a': a->gas = optarg; break;
      case 'C': a->output_type = 0; a->ortho = 0; bad = parse_list(optarg, camera_param, a); break;
      case 'c': a->clouds = optarg; break;
      case 'D': {
        double d[2];
        bad = parse_doubles(optarg, d, 2) || d[0] < 0 || d[0] >= 360 || d[1] < 0 || d[1] > 90;
        a->sun_az = d[0]; a->sun_el = d[1];
      } break;
      case 'd': a->dump = 1; break;
      case 'f': a->force = 1; break;
      case 'g': a->ground = optarg; break;
      case 'h': usage(); return -1;
      case 'i': bad = parse_list(optarg, image_param, a); break;
      case 'M': a->mtl = optarg; break;
      case 'm': a->mie = optarg; break;
      case 'n': a->sky_mtl = optarg; break;
      case 'O': a->cache = optarg; break;
      case 'o': a->output = optarg; break;
      case 'P': a->output_type = 0; a->ortho = 1; bad = parse_list(optarg, ortho_param, a); break;
      case 'p': a->output_type = 1; bad = parse_list(optarg, rect_param, a); break;
      case 'R': a->repeat_ground = 1; break;
      case 'r': a->repeat_clouds = 1; break;
      case 's': bad = parse_list(optarg, spectral_param, a); break;
      case 'T': bad = parse_doubles(optarg, &a->thickness, 1) || a->thickness < 0; break;
      case 't': break; /* thread hint: the GPU path has no host threads */
      case 'V': bad = parse_grid_definition(optarg, a->vdef); break;
      case 'v': a->verbose = 1; break;
      default: usage(); return HTGX_BAD_ARG;
    }
    if(bad) {
      fprintf(stderr, "%s: invalid option argument '%s' -- '%c'\n", argv[0], optarg, opt);
      usage();
      return HTGX_BAD_ARG;
    }
  }
  if(!a->gas) {
    fprintf(stderr, "missing the path of the gas optical properties file -- option '-a'\n");
    return HTGX_BAD_ARG;
  }
  if(a->ground && strcmp(a->ground, "synth") && strcmp(a->ground, "plane") && !a->mtl) {
    fprintf(stderr, "missing the path of the file listing the ground materials -- option '-M'\n");
    return HTGX_BAD_ARG;
  }
  if(a->clouds && strcmp(a->clouds, "synth") && !a->mie) {
    fprintf(stderr, "missing the path toward the file of the Mie's data -- option '-m'\n");
    return HTGX_BAD_ARG;
  }
  if(a->tref <= 0) a->tref = a->spectral == HTGX_SPECTRAL_LW ? 290.0 : 5778.0;
  return HTGX_OK;
}

/* htrdr_accum_get_estimation (src/core/htrdr_accum.h:48-68) */
static void
estimate(double s, double s2, size_t n, double* E, double* SE)
{
  if(!n) { *E = 0; *SE = 0; return; }
  *E = s / (double)n;
  {
    const double V = s2 / (double)n - *E * *E;
    *SE = sqrt((V > 0 ? V : 0) / (double)n);
  }
}

/* dump_buffer (src/atmosphere/htrdr_atmosphere_draw_map.c:477-587) */
static void
dump(FILE* out, const struct args* a, const void* pixels, unsigned w, unsigned h,
     struct htgx_accum* time_acc, struct htgx_accum* flux_acc)
{
  unsigned x, y;
  fprintf(out, "%u %u\n", w, h);
  for(y = 0; y < h; ++y) {
    for(x = 0; x < w; ++x) {
      const struct htgx_accum* t;
      double E, SE;
      if(a->output_type == 1) {
        const struct htgx_pixel_flux* p = (const struct htgx_pixel_flux*)pixels + (size_t)y * w + x;
        if(p->flux.nweights) {
          estimate(p->flux.sum_weights, p->flux.sum_weights_sqr, p->flux.nweights, &E, &SE);
          fprintf(out, "%g %g ", E, SE);
          flux_acc->sum_weights += p->flux.sum_weights;
          flux_acc->sum_weights_sqr += p->flux.sum_weights_sqr;
          flux_acc->nweights += p->flux.nweights;
        } else {
          fprintf(out, "0 0 ");
        }
        fprintf(out, "0 0 0 0 ");
        t = &p->time;
      } else if(a->spectral == HTGX_SPECTRAL_SW_CIE_XYZ) {
        /* dump_pixel_image, ..._draw_map.c:516-527 */
        const struct htgx_pixel_image* p = (const struct htgx_pixel_image*)pixels + (size_t)y * w + x;
        fprintf(out, "%g %g ", p->X.E, p->X.SE);
        fprintf(out, "%g %g ", p->Y.E, p->Y.SE);
        fprintf(out, "%g %g ", p->Z.E, p->Z.SE);
        t = &p->time;
      } else {
        const struct htgx_pixel_xwave* p = (const struct htgx_pixel_xwave*)pixels + (size_t)y * w + x;
        fprintf(out, "%g %g %f %f 0 0 ", p->radiance_temperature.E, p->radiance_temperature.SE,
                p->radiance.E, p->radiance.SE);
        t = &p->time;
      }
      if(t->nweights) {
        estimate(t->sum_weights, t->sum_weights_sqr, t->nweights, &E, &SE);
        fprintf(out, "%g %g ", E, SE);
        time_acc->sum_weights += t->sum_weights;
        time_acc->sum_weights_sqr += t->sum_weights_sqr;
        time_acc->nweights += t->nweights;
      } else {
        fprintf(out, "0 0 ");
      }
      fprintf(out, "\n");
    }
    fprintf(out, "\n");
  }
}

static int
synth_config(const char* spec, int* config, unsigned long long* seed)
{
  char* end;
  if(strncmp(spec, "synth:", 6)) return -1;
  *config = (int)strtol(spec + 6, &end, 10);
  *seed = 0;
  if(*end == ',') *seed = strtoull(end + 1, &end, 10);
  return *end ? -1 : 0;
}

int
main(int argc, char** argv)
{
  struct args a;
  struct htgx_synth_params sp;
  struct htgx_synth* synth = NULL;
  struct htgx_ctx* ctx = NULL;
  struct htgx_ground_files* ground = NULL;
  struct htgx_map_desc d;
  struct htgx_scene_desc scene;
  struct htgx_accum tacc = {0, 0, 0, 0}, facc = {0, 0, 0, 0};
  const char* solver = getenv("HTGX_SOLVER");
  void* pixels = NULL;
  FILE* out = stdout;
  int config, res;
  unsigned long long seed;
  size_t elmt;
  struct timespec t0, t1;

  res = parse_args(argc, argv, &a);
  if(res < 0) return 0; /* -h */
  if(res) return 1;
  if(synth_config(a.gas, &config, &seed)) {
    fprintf(stderr, "%s: `%s': htgop files cannot be read on this path; use synth:<config>[,<seed>]\n",
            argv[0], a.gas);
    return 1;
  }
  if(a.clouds && strcmp(a.clouds, "synth")) {
    fprintf(stderr, "%s: `%s': htcp files cannot be read on this path; use synth\n", argv[0], a.clouds);
    return 1;
  }
  if(a.ground && strcmp(a.ground, "synth") && strcmp(a.ground, "plane")) {
    if(htgx_ground_files_load(a.ground, a.mtl, a.sky_mtl ? a.sky_mtl : "air", &ground) != HTGX_OK) {
      fprintf(stderr, "%s: %s\n", argv[0], htgx_files_last_error());
      return 1;
    }
  }
  if(htgx_synth_params_preset(config, &sp) != HTGX_OK) {
    fprintf(stderr, "%s: unknown synthetic configuration %d\n", argv[0], config);
    return 1;
  }
  sp.seed = seed;
  sp.sun_azimuth = a.sun_az;
  sp.sun_elevation = a.sun_el;
  sp.optical_thickness = a.thickness;
  if(!a.clouds) sp.cloud_def[0] = sp.cloud_def[1] = sp.cloud_def[2] = 0;
  if(!a.ground || ground) sp.ground_type = HTGX_GROUND_NONE;
  else if(!strcmp(a.ground, "plane")) sp.ground_type = HTGX_GROUND_PLANE;
  else {
    if(sp.ground_type != HTGX_GROUND_MESH) {
      sp.hf_res = 257; sp.hf_octaves = 8; sp.hf_relief = 800.0; sp.hf_hurst = 0.7;
    }
    sp.ground_type = HTGX_GROUND_MESH;
  }
  if((res = htgx_synth_create(&sp, &synth)) != HTGX_OK) {
    fprintf(stderr, "%s: could not create the synthetic atmosphere (%d)\n", argv[0], res);
    htgx_ground_files_destroy(ground);
    return 1;
  }

  memset(&d, 0, sizeof(d));
  d.width = a.def[0];
  d.height = a.def[1];
  d.spp = a.spp;
  d.spectral_type = a.spectral;
  d.wlen_range[0] = a.wlen[0];
  d.wlen_range[1] = a.wlen[1];
  d.ref_temperature = a.tref;
  d.solver = solver && !strcmp(solver, "plain") ? HTGX_SOLVER_PLAIN_FP64 : HTGX_SOLVER_WAVEFRONT_FP32;
  if(a.output_type == 1) {
    d.sensor = HTGX_SENSOR_RECTANGLE;
    memcpy(d.pos, a.rpos, sizeof(d.pos));
    memcpy(d.tgt, a.rtgt, sizeof(d.tgt));
    memcpy(d.up, a.rup, sizeof(d.up));
    d.rect_size[0] = a.rsz[0];
    d.rect_size[1] = a.rsz[1];
    elmt = sizeof(struct htgx_pixel_flux);
  } else {
    d.sensor = HTGX_SENSOR_CAMERA;
    memcpy(d.pos, a.pos, sizeof(d.pos));
    memcpy(d.tgt, a.tgt, sizeof(d.tgt));
    memcpy(d.up, a.up, sizeof(d.up));
    d.fov_y = a.fov;
    d.projection = a.ortho ? HTGX_PROJECTION_ORTHOGRAPHIC : HTGX_PROJECTION_PERSPECTIVE;
    d.lens_radius = a.lens_radius;
    d.focal_dst = a.focal_dst;
    d.ortho_height = a.ortho_height;
    elmt = a.spectral == HTGX_SPECTRAL_SW_CIE_XYZ
      ? sizeof(struct htgx_pixel_image) : sizeof(struct htgx_pixel_xwave);
  }

  if(a.output) {
    const int fd = open(a.output, O_WRONLY | O_CREAT | (a.force ? O_TRUNC : O_EXCL), 0644);
    if(fd < 0 || !(out = fdopen(fd, "w"))) {
      fprintf(stderr, "%s: cannot open `%s' for writing%s\n", argv[0], a.output,
              a.force ? "" : " (use -f to overwrite)");
      htgx_synth_destroy(synth);
      htgx_ground_files_destroy(ground);
      return 1;
    }
  }

  scene = *htgx_synth_get_scene(synth);
  if(ground) scene.ground = *htgx_ground_files_get_desc(ground);
  scene.repeat_clouds = a.repeat_clouds;
  scene.repeat_ground = a.repeat_ground;
  if(a.vdef[0] && scene.clouds.def[0]) { /* -V: the octree's finest definition */
    uint32_t B = 0;
    const uint32_t vdef[3] = {a.vdef[0], a.vdef[1], a.vdef[2]};
    if(htgx_leaf_block_for_definition(scene.clouds.def, vdef, &B) != HTGX_OK) {
      fprintf(stderr, "%s: no block of cells divides the cloud grid %ux%ux%u into at most "
              "%u,%u,%u blocks -- option '-V'\n", argv[0], scene.clouds.def[0], scene.clouds.def[1],
              scene.clouds.def[2], vdef[0], vdef[1], vdef[2]);
      res = HTGX_BAD_ARG;
      goto error;
    }
    scene.leaf_block = B;
  }
  if(a.dump) { /* dump_volumetric_acceleration_structure, ..._atmosphere.c:228-255 */
    fflush(out);
    res = htgx_dump_octrees_vtk(&scene, fileno(out));
    if(res == HTGX_BAD_OP) fprintf(stderr, "%s: no clouds, no octree to write (-d)\n", argv[0]);
    else if(res != HTGX_OK) fprintf(stderr, "%s: could not write the octrees (%d)\n", argv[0], res);
    goto error;
  }
  if(a.cache) { /* htsky's cache_filename, src/atmosphere/htrdr_atmosphere.c:369 */
    char msg[512];
    int loaded = 0;
    if((res = htgx_ctx_create_cached(&scene, a.cache, &loaded, msg, sizeof(msg), &ctx)) != HTGX_OK) {
      fprintf(stderr, "%s: could not set up the scene on the GPU (%d)%s%s\n", argv[0], res,
              msg[0] ? ": " : "", msg);
      goto error;
    }
  } else if((res = htgx_ctx_create(&scene, &ctx)) != HTGX_OK) {
    fprintf(stderr, "%s: could not set up the scene on the GPU (%d)\n", argv[0], res);
    goto error;
  }
  pixels = calloc((size_t)d.width * d.height, elmt);
  if(!pixels) { res = HTGX_MEM_ERR; goto error; }
  clock_gettime(CLOCK_MONOTONIC, &t0);
  res = htgx_draw_map(ctx, &d, pixels, (size_t)d.width * elmt);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if(res != HTGX_OK) {
    fprintf(stderr, "%s: rendering failed (%d)\n", argv[0], res);
    goto error;
  }
  if(a.verbose) {
    const double s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
    fprintf(stderr, "Rendering time: %.3f s (%.3f Mrealisations/s)\n", s,
            (double)d.width * d.height * d.spp * (a.spectral == HTGX_SPECTRAL_SW_CIE_XYZ ? 3 : 1)
              / s * 1e-6);
  }
  dump(out, &a, pixels, d.width, d.height, &tacc, &facc);
  {
    double E, SE;
    estimate(tacc.sum_weights, tacc.sum_weights_sqr, tacc.nweights, &E, &SE);
    fprintf(stderr, "Time per radiative path (in micro seconds): %g +/- %g\n", E, SE);
    if(a.output_type == 1) {
      estimate(facc.sum_weights, facc.sum_weights_sqr, facc.nweights, &E, &SE);
      fprintf(stderr, "Radiative flux density (in W/(external m^2)): %g +/- %g\n", E, SE);
    }
  }
error:
  free(pixels);
  if(ctx) htgx_ctx_destroy(ctx);
  htgx_synth_destroy(synth);
  htgx_ground_files_destroy(ground);
  if(out != stdout) fclose(out);
  return res == HTGX_OK ? 0 : 1;
}
