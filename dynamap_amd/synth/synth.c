/*
 * Seeded synthetic RGB-D generator for the rigid-fusion workloads.
 *
 * Scenes are analytic: the interior of an axis-aligned room, solid boxes and
 * spheres.  A frame is rendered by exact ray casting in FP64 (pinhole camera,
 * z-depth), then Kinect-like axial noise N(0, 0.0012 + 0.0019 (z - 0.4)^2) and
 * 1 % dropout (depth = 0) are applied.  Colour is a seeded value-noise albedo
 * evaluated at the hit point in world coordinates, so it is view independent
 * (Lambertian, no shading) and a photometric residual is consistent across
 * frames.  Trajectories are a smooth random walk clamped to 2 cm and 1 degree
 * per frame.
 *
 * All randomness comes from SplitMix64 on (seed, stream, counter), so every
 * output is a pure function of its arguments and independent of threading.
 * This file is harness code for tests and benchmarks, not part of the mapping
 * path.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define DMS_MAX_PRIMS 64

enum { PRIM_ROOM = 0, PRIM_BOX = 1, PRIM_SPHERE = 2 };

typedef struct {
  int kind;
  double lo[3], hi[3]; /* boxes and room */
  double c[3], r;      /* spheres */
  double rgb[3];       /* base albedo in [0,1] */
} dms_prim;

typedef struct {
  int nprims;
  dms_prim prims[DMS_MAX_PRIMS];
  double room_lo[3], room_hi[3];
  double cam_center[3]; /* trajectories stay near this point */
  double cam_radius;    /* ... within this radius */
  uint64_t seed;
} dms_scene;

static uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

static uint64_t hash3(uint64_t seed, uint64_t a, uint64_t b) {
  return splitmix64(splitmix64(splitmix64(seed) ^ a) ^ b);
}

/* uniform in [0,1) with 53 random bits */
static double u01(uint64_t h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

typedef struct { uint64_t seed, ctr; } dms_rng;
static double rng_u(dms_rng *r) { return u01(hash3(r->seed, 0x5EEDULL, r->ctr++)); }
static double rng_range(dms_rng *r, double a, double b) { return a + (b - a) * rng_u(r); }
static double rng_normal(dms_rng *r) {
  double u1 = rng_u(r), u2 = rng_u(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static void add_sphere(dms_scene *s, double x, double y, double z, double r, dms_rng *g) {
  dms_prim *p = &s->prims[s->nprims++];
  memset(p, 0, sizeof(*p));
  p->kind = PRIM_SPHERE;
  p->c[0] = x; p->c[1] = y; p->c[2] = z; p->r = r;
  for (int k = 0; k < 3; ++k) p->rgb[k] = rng_range(g, 0.25, 1.0);
}

static void add_box(dms_scene *s, const double lo[3], const double hi[3], int kind, dms_rng *g) {
  dms_prim *p = &s->prims[s->nprims++];
  memset(p, 0, sizeof(*p));
  p->kind = kind;
  for (int k = 0; k < 3; ++k) { p->lo[k] = lo[k]; p->hi[k] = hi[k]; p->rgb[k] = rng_range(g, 0.25, 1.0); }
}

static double dist_point_box(const double q[3], const double lo[3], const double hi[3]) {
  double d2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double e = q[k] < lo[k] ? lo[k] - q[k] : (q[k] > hi[k] ? q[k] - hi[k] : 0.0);
    d2 += e * e;
  }
  return sqrt(d2);
}

/*
 * kind 1: 4 x 3 x 3 m room with 5 spheres, r ~ U[0.2, 0.5].
 * kind 5: 8 x 8 x 3 m multi-room floor with 12 boxes (4 of them wall
 *         segments splitting the floor into rooms) and 20 spheres.
 * Objects keep clear of a ball around cam_center where the camera moves.
 */
int dms_scene_init(dms_scene *s, int kind, uint64_t seed) {
  memset(s, 0, sizeof(*s));
  s->seed = seed;
  dms_rng g = {seed, 0};
  if (kind == 1) {
    double lo[3] = {-2.0, -1.5, -1.5}, hi[3] = {2.0, 1.5, 1.5};
    memcpy(s->room_lo, lo, sizeof lo); memcpy(s->room_hi, hi, sizeof hi);
    add_box(s, lo, hi, PRIM_ROOM, &g);
    s->cam_center[0] = 0.0; s->cam_center[1] = 0.0; s->cam_center[2] = 0.0;
    s->cam_radius = 0.35;
    int placed = 0, tries = 0;
    while (placed < 5 && tries < 100000) {
      ++tries;
      double r = rng_range(&g, 0.2, 0.5);
      double c[3];
      for (int k = 0; k < 3; ++k) c[k] = rng_range(&g, lo[k] + r, hi[k] - r);
      c[1] = rng_range(&g, hi[1] - r - 0.5, hi[1] - r); /* resting near the floor (+y is down) */
      double dc = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      if (dc < r + s->cam_radius + 0.45) continue; /* keep the camera ball clear */
      int ok = 1;
      for (int i = 1; i < s->nprims; ++i) {
        const dms_prim *o = &s->prims[i];
        double dx = o->c[0] - c[0], dy = o->c[1] - c[1], dz = o->c[2] - c[2];
        if (sqrt(dx * dx + dy * dy + dz * dz) < o->r + r + 0.05) ok = 0;
      }
      if (!ok) continue;
      add_sphere(s, c[0], c[1], c[2], r, &g);
      ++placed;
    }
    return placed == 5 ? 0 : -1;
  }
  if (kind == 5) {
    double lo[3] = {-4.0, -1.5, -4.0}, hi[3] = {4.0, 1.5, 4.0};
    memcpy(s->room_lo, lo, sizeof lo); memcpy(s->room_hi, hi, sizeof hi);
    add_box(s, lo, hi, PRIM_ROOM, &g);
    s->cam_center[0] = -1.2; s->cam_center[1] = 0.0; s->cam_center[2] = -1.2;
    s->cam_radius = 0.5;
    /* four wall segments with door gaps: rooms around the origin */
    double w[4][6] = {
        {-4.0, -1.5, 0.9, -1.9, 1.5, 1.05}, {-0.6, -1.5, 0.9, 4.0, 1.5, 1.05},
        {0.9, -1.5, -4.0, 1.05, 1.5, -2.1}, {0.9, -1.5, -0.2, 1.05, 1.5, 0.9}};
    for (int i = 0; i < 4; ++i) add_box(s, w[i], w[i] + 3, PRIM_BOX, &g);
    int boxes = 0, tries = 0;
    while (boxes < 8 && tries < 100000) {
      ++tries;
      double e[3] = {rng_range(&g, 0.3, 0.9), rng_range(&g, 0.3, 1.2), rng_range(&g, 0.3, 0.9)};
      double c[3] = {rng_range(&g, lo[0] + e[0], hi[0] - e[0]), hi[1] - e[1],
                     rng_range(&g, lo[2] + e[2], hi[2] - e[2])};
      double blo[3] = {c[0] - e[0], c[1] - e[1], c[2] - e[2]};
      double bhi[3] = {c[0] + e[0], c[1] + e[1], c[2] + e[2]};
      if (dist_point_box(s->cam_center, blo, bhi) < s->cam_radius + 0.6) continue;
      add_box(s, blo, bhi, PRIM_BOX, &g);
      ++boxes;
    }
    int sph = 0;
    tries = 0;
    while (sph < 20 && tries < 100000) {
      ++tries;
      double r = rng_range(&g, 0.15, 0.45);
      double c[3];
      for (int k = 0; k < 3; ++k) c[k] = rng_range(&g, lo[k] + r, hi[k] - r);
      double dx = c[0] - s->cam_center[0], dy = c[1] - s->cam_center[1], dz = c[2] - s->cam_center[2];
      if (sqrt(dx * dx + dy * dy + dz * dz) < r + s->cam_radius + 0.6) continue;
      add_sphere(s, c[0], c[1], c[2], r, &g);
      ++sph;
    }
    return (boxes == 8 && sph == 20) ? 0 : -1;
  }
  return -2;
}

/* Value noise in [0,1] on a lattice of `cell` metres. */
static double value_noise(uint64_t seed, const double x[3], double cell) {
  double q[3];
  int64_t i0[3];
  double f[3];
  for (int k = 0; k < 3; ++k) {
    q[k] = x[k] / cell;
    double fl = floor(q[k]);
    i0[k] = (int64_t)fl;
    double t = q[k] - fl;
    f[k] = t * t * (3.0 - 2.0 * t);
  }
  double acc = 0.0;
  for (int c = 0; c < 8; ++c) {
    int dx = c & 1, dy = (c >> 1) & 1, dz = (c >> 2) & 1;
    uint64_t h = hash3(seed, (uint64_t)(i0[0] + dx) * 0x100000001B3ULL ^ (uint64_t)(i0[1] + dy),
                       (uint64_t)(i0[2] + dz));
    double w = (dx ? f[0] : 1.0 - f[0]) * (dy ? f[1] : 1.0 - f[1]) * (dz ? f[2] : 1.0 - f[2]);
    acc += w * u01(h);
  }
  return acc;
}

/* Nearest hit along o + t d (t > 0); returns t and the primitive index. */
static double cast(const dms_scene *s, const double o[3], const double d[3], int *hit_prim) {
  double best = INFINITY;
  int bp = -1;
  for (int i = 0; i < s->nprims; ++i) {
    const dms_prim *p = &s->prims[i];
    if (p->kind == PRIM_SPHERE) {
      double oc[3] = {o[0] - p->c[0], o[1] - p->c[1], o[2] - p->c[2]};
      double a = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
      double b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
      double c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - p->r * p->r;
      double disc = b * b - a * c;
      if (disc < 0.0) continue;
      double sq = sqrt(disc);
      double t = (-b - sq) / a;
      if (t <= 1e-9) t = (-b + sq) / a;
      if (t > 1e-9 && t < best) { best = t; bp = i; }
    } else {
      double tmin = -INFINITY, tmax = INFINITY;
      int ok = 1;
      for (int k = 0; k < 3; ++k) {
        if (d[k] == 0.0) {
          if (o[k] < p->lo[k] || o[k] > p->hi[k]) ok = 0;
          continue;
        }
        double t1 = (p->lo[k] - o[k]) / d[k], t2 = (p->hi[k] - o[k]) / d[k];
        if (t1 > t2) { double tt = t1; t1 = t2; t2 = tt; }
        if (t1 > tmin) tmin = t1;
        if (t2 < tmax) tmax = t2;
      }
      if (!ok || tmin > tmax) continue;
      double t = (p->kind == PRIM_ROOM) ? tmax : (tmin > 1e-9 ? tmin : INFINITY);
      if (t > 1e-9 && t < best) { best = t; bp = i; }
    }
  }
  *hit_prim = bp;
  return best;
}

/*
 * Render one frame.  pose: row-major 4x4 camera-to-world (x right, y down,
 * z forward).  depth: rows*cols float32 metres (0 = no return).  rgb: rows*cols*3
 * uint8 or NULL.  noise != 0 applies axial noise and 1 % dropout.
 */
void dms_render(const dms_scene *s, const double pose[16], int rows, int cols, double fx,
                double fy, double cx, double cy, uint64_t frame_seed, int noise, float *depth,
                uint8_t *rgb) {
  const double o[3] = {pose[3], pose[7], pose[11]};
#pragma omp parallel for schedule(static)
  for (int v = 0; v < rows; ++v) {
    for (int u = 0; u < cols; ++u) {
      int i = v * cols + u;
      double dc[3] = {((double)u - cx) / fx, ((double)v - cy) / fy, 1.0};
      double d[3];
      for (int r = 0; r < 3; ++r) d[r] = pose[4 * r + 0] * dc[0] + pose[4 * r + 1] * dc[1] + pose[4 * r + 2] * dc[2];
      int prim;
      double t = cast(s, o, d, &prim);
      double z = (prim >= 0 && isfinite(t)) ? t : 0.0;
      if (rgb) {
        uint8_t c3[3] = {0, 0, 0};
        if (prim >= 0) {
          double x[3] = {o[0] + t * d[0], o[1] + t * d[1], o[2] + t * d[2]};
          double n = value_noise(s->seed ^ 0xA1BED0ULL, x, 0.07);
          double n2 = value_noise(s->seed ^ 0xC0102ULL, x, 0.23);
          for (int k = 0; k < 3; ++k) {
            double val = s->prims[prim].rgb[k] * (0.25 + 0.5 * n + 0.25 * n2);
            int q = (int)floor(val * 255.0 + 0.5);
            c3[k] = (uint8_t)(q < 0 ? 0 : (q > 255 ? 255 : q));
          }
        }
        rgb[3 * i + 0] = c3[0]; rgb[3 * i + 1] = c3[1]; rgb[3 * i + 2] = c3[2];
      }
      if (noise && z > 0.0) {
        dms_rng g = {hash3(frame_seed, 0xD5EULL, (uint64_t)i), 0};
        if (rng_u(&g) < 0.01) {
          z = 0.0;
        } else {
          double sig = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
          z += sig * rng_normal(&g);
          if (z < 0.0) z = 0.0;
        }
      }
      depth[i] = (float)z;
    }
  }
}

/* Rotation matrix (row-major) from axis-angle w (radians). */
static void rodrigues(const double w[3], double R[9]) {
  double th = sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
  double k[3] = {0, 0, 0};
  if (th > 0) { k[0] = w[0] / th; k[1] = w[1] / th; k[2] = w[2] / th; }
  double c = cos(th), sn = sin(th), C = 1.0 - c;
  R[0] = c + k[0] * k[0] * C;        R[1] = k[0] * k[1] * C - k[2] * sn; R[2] = k[0] * k[2] * C + k[1] * sn;
  R[3] = k[1] * k[0] * C + k[2] * sn; R[4] = c + k[1] * k[1] * C;        R[5] = k[1] * k[2] * C - k[0] * sn;
  R[6] = k[2] * k[0] * C - k[1] * sn; R[7] = k[2] * k[1] * C + k[0] * sn; R[8] = c + k[2] * k[2] * C;
}

static void matmul3(const double A[9], const double B[9], double C[9]) {
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) C[3 * r + c] = A[3 * r] * B[c] + A[3 * r + 1] * B[3 + c] + A[3 * r + 2] * B[6 + c];
}

/* camera-to-world rotation (columns = camera x right, y down, z forward)
 * looking along f with the camera y axis close to world +y (down) */
static void look_rotation(const double f_in[3], double R[9]) {
  double fn = sqrt(f_in[0] * f_in[0] + f_in[1] * f_in[1] + f_in[2] * f_in[2]);
  double f[3] = {f_in[0] / fn, f_in[1] / fn, f_in[2] / fn};
  double down[3] = {0.0, 1.0, 0.0};
  double x[3] = {down[1] * f[2] - down[2] * f[1], down[2] * f[0] - down[0] * f[2], down[0] * f[1] - down[1] * f[0]};
  double xn = sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  for (int k = 0; k < 3; ++k) x[k] /= xn;
  double y[3] = {f[1] * x[2] - f[2] * x[1], f[2] * x[0] - f[0] * x[2], f[0] * x[1] - f[1] * x[0]};
  for (int r = 0; r < 3; ++r) { R[3 * r] = x[r]; R[3 * r + 1] = y[r]; R[3 * r + 2] = f[r]; }
}

/* axis-angle of the rotation A^T B (angle in [0, pi]) */
static double rel_axis_angle(const double A[9], const double B[9], double w[3]) {
  double M[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) M[3 * r + c] = A[r] * B[c] + A[3 + r] * B[3 + c] + A[6 + r] * B[6 + c];
  double cs = (M[0] + M[4] + M[8] - 1.0) * 0.5;
  if (cs > 1.0) cs = 1.0;
  if (cs < -1.0) cs = -1.0;
  double th = acos(cs);
  double v[3] = {M[7] - M[5], M[2] - M[6], M[3] - M[1]};
  double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  for (int k = 0; k < 3; ++k) w[k] = vn > 1e-12 ? v[k] / vn * th : 0.0;
  return th;
}

/*
 * Smooth seeded random walk: per-frame translation <= 2 cm and rotation
 * <= 1 degree, staying within cam_radius of cam_center.  The camera looks at
 * the scene's objects: its heading follows a look-at target (the centroid of
 * the spheres, swept +-25 degrees in yaw), approached at most 1 degree per
 * frame including a small random jitter, so that the view keeps enough
 * geometry to constrain tracking.  poses: nframes * 16 row-major
 * camera-to-world.
 */
void dms_trajectory(const dms_scene *s, int nframes, uint64_t seed, double *poses) {
  dms_rng g = {seed ^ 0x7A3EC7ULL, 0};
  const double PI = 3.141592653589793;
  const double max_step = 0.02, max_rot = 1.0 * PI / 180.0;
  double p[3] = {s->cam_center[0], s->cam_center[1], s->cam_center[2]};
  double tgt[3] = {0, 0, 0};
  int ns = 0;
  for (int i = 0; i < s->nprims; ++i)
    if (s->prims[i].kind == PRIM_SPHERE) { for (int k = 0; k < 3; ++k) tgt[k] += s->prims[i].c[k]; ++ns; }
  if (ns) for (int k = 0; k < 3; ++k) tgt[k] /= ns;
  const double phase = rng_range(&g, 0.0, 2.0 * PI);
  double v[3] = {0, 0, 0};
  double R[9];
  for (int f = 0; f < nframes; ++f) {
    if (f > 0) {
      for (int k = 0; k < 3; ++k) {
        double pull = (s->cam_center[k] - p[k]) * 0.02;
        v[k] = 0.9 * v[k] + 0.003 * rng_normal(&g) + pull;
      }
      double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      if (vn > max_step) for (int k = 0; k < 3; ++k) v[k] *= max_step / vn;
      for (int k = 0; k < 3; ++k) p[k] += v[k];
      double dx = p[0] - s->cam_center[0], dy = p[1] - s->cam_center[1], dz = p[2] - s->cam_center[2];
      double dn = sqrt(dx * dx + dy * dy + dz * dz);
      if (dn > s->cam_radius) {
        p[0] = s->cam_center[0] + dx * s->cam_radius / dn;
        p[1] = s->cam_center[1] + dy * s->cam_radius / dn;
        p[2] = s->cam_center[2] + dz * s->cam_radius / dn;
      }
    }
    /* desired heading: towards the target, yaw swept, pitched 15..35 deg down */
    double d[3] = {tgt[0] - p[0], tgt[1] - p[1], tgt[2] - p[2]};
    double yaw = atan2(d[0], d[2]) + 25.0 * PI / 180.0 * sin(2.0 * PI * f / 150.0 + phase);
    double hor = sqrt(d[0] * d[0] + d[2] * d[2]);
    double pitch = atan2(d[1], hor);
    if (pitch < 15.0 * PI / 180.0) pitch = 15.0 * PI / 180.0;
    if (pitch > 35.0 * PI / 180.0) pitch = 35.0 * PI / 180.0;
    double fw[3] = {sin(yaw) * cos(pitch), sin(pitch), cos(yaw) * cos(pitch)};
    double Rd[9];
    look_rotation(fw, Rd);
    if (f == 0) {
      memcpy(R, Rd, sizeof R);
    } else {
      /* rotate towards Rd by at most 0.8 deg, plus up to 0.2 deg of jitter */
      double w[3];
      double th = rel_axis_angle(R, Rd, w);
      double lim = 0.8 * PI / 180.0;
      if (th > lim) for (int k = 0; k < 3; ++k) w[k] *= lim / th;
      double jit[3] = {rng_normal(&g), rng_normal(&g), rng_normal(&g)};
      double jn = sqrt(jit[0] * jit[0] + jit[1] * jit[1] + jit[2] * jit[2]);
      for (int k = 0; k < 3; ++k) w[k] += jit[k] / (jn > 0 ? jn : 1.0) * (max_rot - lim) * rng_u(&g);
      double dR[9], Rn[9];
      rodrigues(w, dR);  /* w is in the camera frame: R <- R dR */
      matmul3(R, dR, Rn);
      memcpy(R, Rn, sizeof R);
    }
    double *P = poses + 16 * (size_t)f;
    P[0] = R[0]; P[1] = R[1]; P[2] = R[2];  P[3] = p[0];
    P[4] = R[3]; P[5] = R[4]; P[6] = R[5];  P[7] = p[1];
    P[8] = R[6]; P[9] = R[7]; P[10] = R[8]; P[11] = p[2];
    P[12] = 0;   P[13] = 0;   P[14] = 0;    P[15] = 1;
  }
}

int dms_scene_size(void) { return (int)sizeof(dms_scene); }

/* ------------------------------------------------------------------ */
/* Deforming object: a bending cylinder                                 */
/* ------------------------------------------------------------------ */
/*
 * Canonical shape: lateral surface of a cylinder of radius 0.15 m, axis
 * along world y (down), y in [-0.3, 0.3], centred at the origin, plus the
 * top cap at y = -0.3.  Deformation at frame f: every point moves along x
 * by b(y, f) = A sin(2 pi f / 40) s(y)^2 with A = 0.05 m and
 * s(y) = (0.3 - y) / 0.6 (0 at the bottom, which stays fixed, 1 at the top).
 */
#define CYL_R 0.15
#define CYL_H2 0.3
#define CYL_A 0.05
#define CYL_PERIOD 40.0

static double cyl_bend(double y, double f) {
  double s = (CYL_H2 - y) / (2.0 * CYL_H2);
  return CYL_A * sin(2.0 * 3.141592653589793 * f / CYL_PERIOD) * s * s;
}

/* Ground-truth deformation of n canonical points (xyz triples, FP64). */
void dms_cylinder_warp(double frame, int64_t n, const double *in, double *out) {
  for (int64_t i = 0; i < n; ++i) {
    out[3 * i] = in[3 * i] + cyl_bend(in[3 * i + 1], frame);
    out[3 * i + 1] = in[3 * i + 1];
    out[3 * i + 2] = in[3 * i + 2];
  }
}

/* Canonical mesh: n_theta x n_y vertices of the lateral surface (index
 * j * n_theta + i, theta_i = 2 pi i / n_theta, y_j evenly from -0.3 to 0.3)
 * with outward normals, float32. */
void dms_cylinder_mesh(int n_theta, int n_y, float *verts, float *normals) {
  for (int j = 0; j < n_y; ++j) {
    double y = -CYL_H2 + 2.0 * CYL_H2 * (double)j / (double)(n_y - 1);
    for (int i = 0; i < n_theta; ++i) {
      double th = 2.0 * 3.141592653589793 * (double)i / (double)n_theta;
      size_t k = (size_t)j * n_theta + i;
      verts[3 * k] = (float)(CYL_R * cos(th));
      verts[3 * k + 1] = (float)y;
      verts[3 * k + 2] = (float)(CYL_R * sin(th));
      normals[3 * k] = (float)cos(th);
      normals[3 * k + 1] = 0.f;
      normals[3 * k + 2] = (float)sin(th);
    }
  }
}

/* first hit of the ray o + t d (t > 0) with the deformed cylinder, or -1 */
static double cyl_cast(const double o[3], const double d[3], double f) {
  double best = -1.0;
  /* top cap: plane y = -0.3, disc around (b(-0.3), 0) */
  if (fabs(d[1]) > 1e-12) {
    double t = (-CYL_H2 - o[1]) / d[1];
    if (t > 0) {
      double x = o[0] + t * d[0] - cyl_bend(-CYL_H2, f), z = o[2] + t * d[2];
      if (x * x + z * z <= CYL_R * CYL_R) best = t;
    }
  }
  /* lateral surface: g(t) = (x - b(y))^2 + z^2 - r^2, marched in 2 mm steps
   * over the slab |y| <= 0.3 and the bounding box, then bisected */
  double t0 = 1e-6, t1 = 20.0;
  const double lo[3] = {-CYL_R - CYL_A, -CYL_H2, -CYL_R}, hi[3] = {CYL_R + CYL_A, CYL_H2, CYL_R};
  for (int k = 0; k < 3; ++k) {
    if (fabs(d[k]) < 1e-12) {
      if (o[k] < lo[k] || o[k] > hi[k]) return best;
      continue;
    }
    double a = (lo[k] - o[k]) / d[k], b = (hi[k] - o[k]) / d[k];
    if (a > b) { double tmp = a; a = b; b = tmp; }
    if (a > t0) t0 = a;
    if (b < t1) t1 = b;
  }
  if (t0 >= t1) return best;
  const double step = 0.002;
#define CYL_G(t) ({ double x_ = o[0] + (t) * d[0], y_ = o[1] + (t) * d[1], z_ = o[2] + (t) * d[2]; \
                    double u_ = x_ - cyl_bend(y_, f); u_ * u_ + z_ * z_ - CYL_R * CYL_R; })
  double ta = t0, ga = CYL_G(ta);
  for (double tb = t0 + step;; tb += step) {
    if (tb > t1) tb = t1;
    double gb = CYL_G(tb);
    if (ga > 0 && gb <= 0) {
      double a = ta, b = tb;
      for (int it = 0; it < 60; ++it) {
        double m = 0.5 * (a + b);
        if (CYL_G(m) > 0) a = m; else b = m;
      }
      double t = 0.5 * (a + b);
      double y = o[1] + t * d[1];
      if (y >= -CYL_H2 && y <= CYL_H2 && (best < 0 || t < best)) best = t;
      break;
    }
    if (tb >= t1) break;
    ta = tb;
    ga = gb;
  }
#undef CYL_G
  return best;
}

/* Depth frame of the cylinder deformed at frame f, seen from pose; the same
 * camera model and noise as dms_render. */
void dms_cylinder_render(double f, const double pose[16], int rows, int cols, double fx, double fy, double cx,
                         double cy, uint64_t frame_seed, int noise, float *depth) {
  const double o[3] = {pose[3], pose[7], pose[11]};
#pragma omp parallel for schedule(dynamic, 4)
  for (int v = 0; v < rows; ++v) {
    for (int u = 0; u < cols; ++u) {
      int i = v * cols + u;
      double dc[3] = {((double)u - cx) / fx, ((double)v - cy) / fy, 1.0};
      double d[3];
      for (int r = 0; r < 3; ++r) d[r] = pose[4 * r + 0] * dc[0] + pose[4 * r + 1] * dc[1] + pose[4 * r + 2] * dc[2];
      double z = cyl_cast(o, d, f);
      if (z < 0) z = 0.0;
      if (noise && z > 0.0) {
        dms_rng g = {hash3(frame_seed, 0xD5EULL, (uint64_t)i), 0};
        if (rng_u(&g) < 0.01) {
          z = 0.0;
        } else {
          double sig = 0.0012 + 0.0019 * (z - 0.4) * (z - 0.4);
          z += sig * rng_normal(&g);
          if (z < 0.0) z = 0.0;
        }
      }
      depth[i] = (float)z;
    }
  }
}

/* Camera poses for the cylinder sequence: 0.9 m from the axis, 0.3 m above
 * the centre, looking at the origin, with a seeded drift of at most 5 mm and
 * 0.3 degrees per frame around that nominal pose (bounded to 2 cm). */
void dms_cylinder_trajectory(int nframes, uint64_t seed, double *poses) {
  dms_rng g = {seed ^ 0xC711DEULL, 0};
  double base[3] = {0.0, -0.45, -0.85};
  double p[3] = {base[0], base[1], base[2]};
  double yaw = 0.0, pitch = 0.0;
  for (int f = 0; f < nframes; ++f) {
    if (f > 0) {
      for (int k = 0; k < 3; ++k) {
        double stepk = 0.0025 * rng_normal(&g);
        if (stepk > 0.005) stepk = 0.005;
        if (stepk < -0.005) stepk = -0.005;
        p[k] += stepk - 0.05 * (p[k] - base[k]);
        if (p[k] - base[k] > 0.02) p[k] = base[k] + 0.02;
        if (p[k] - base[k] < -0.02) p[k] = base[k] - 0.02;
      }
      yaw += 0.003 * rng_normal(&g) - 0.05 * yaw;
      pitch += 0.003 * rng_normal(&g) - 0.05 * pitch;
    }
    double tgt[3] = {0.0, 0.0, 0.0};
    double fw0[3] = {tgt[0] - p[0], tgt[1] - p[1], tgt[2] - p[2]};
    double hor = sqrt(fw0[0] * fw0[0] + fw0[2] * fw0[2]);
    double y0 = atan2(fw0[0], fw0[2]) + yaw, p0 = atan2(fw0[1], hor) + pitch;
    double fw[3] = {sin(y0) * cos(p0), sin(p0), cos(y0) * cos(p0)};
    double R[9];
    look_rotation(fw, R);
    double *P = poses + 16 * (size_t)f;
    P[0] = R[0]; P[1] = R[1]; P[2] = R[2];  P[3] = p[0];
    P[4] = R[3]; P[5] = R[4]; P[6] = R[5];  P[7] = p[1];
    P[8] = R[6]; P[9] = R[7]; P[10] = R[8]; P[11] = p[2];
    P[12] = 0;   P[13] = 0;   P[14] = 0;    P[15] = 1;
  }
}
