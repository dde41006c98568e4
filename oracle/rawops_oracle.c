/*
 * CPU oracle for the tiled RAW render path -- TEST INFRASTRUCTURE ONLY.
 * Restates the reference RawEngine arithmetic (tylermcm/librawops) in plain C
 * so that tests can check the HIP kernels.  Never linked into the product.
 * Compile with -ffp-contract=off: the reference's FP64/FP32 operation order is
 * reproduced term by term and must not be contracted into FMAs.
 */
#include "rawops_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ---------------------------------------------------------------- sensor -- */

/* site = row-major index in the pattern-relative 2x2 cell, RawEngine.cpp:74-77. */
static inline unsigned site_of(const ro_bayer* m, int64_t x, int64_t y) {
    return (((unsigned)(y & 1) + m->phase_y) & 1u) * 2u + (((unsigned)(x & 1) + m->phase_x) & 1u);
}

/* Colour of a photosite (0 R, 1 G, 2 B), RawEngine.cpp:79-87. */
static inline int color_of(const ro_bayer* m, int64_t x, int64_t y) {
    static const int cells[4][4] = {{0, 1, 1, 2}, {2, 1, 1, 0}, {1, 0, 2, 1}, {1, 2, 0, 1}};
    return cells[m->pattern][site_of(m, x, y)];
}

/* Signed, unclipped normalisation, RawEngine.cpp:89-97. */
static inline float normalized(const ro_bayer* m, const uint16_t* s, int64_t x, int64_t y) {
    const unsigned site = site_of(m, x, y);
    const uint16_t code = s[(size_t)y * m->row_stride + (size_t)x];
    return ((float)code - (float)m->black[site]) / (float)((int)m->white[site] - (int)m->black[site]);
}

static inline int inside(const ro_rect* a, int64_t x, int64_t y) {
    return x >= a->x && y >= a->y && x < (int64_t)a->x + a->width && y < (int64_t)a->y + a->height;
}

/* ------------------------------------------------------------- bilinear -- */

void ro_demosaic_bilinear(const ro_bayer* m, const uint16_t* s, ro_rect r, float* out) {
    const ro_rect a = m->area;
    for (uint32_t row = 0; row < r.height; ++row) {
        const int64_t y = (int64_t)r.y + row;
        for (uint32_t col = 0; col < r.width; ++col) {
            const int64_t x = (int64_t)r.x + col;
            float* px = out + ((size_t)row * r.width + col) * 3;
            const int own = color_of(m, x, y);
            for (int c = 0; c < 3; ++c) {
                if (own == c) { px[c] = normalized(m, s, x, y); continue; }
                float sum = 0.0f;
                int count = 0;
                for (int dy = -1; dy <= 1; ++dy) {
                    const int64_t yy = y + dy;
                    if (yy < a.y || yy >= (int64_t)a.y + a.height) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int64_t xx = x + dx;
                        if (xx < a.x || xx >= (int64_t)a.x + a.width) continue;
                        if (color_of(m, xx, yy) == c) { sum += normalized(m, s, xx, yy); ++count; }
                    }
                }
                px[c] = count ? sum / (float)count : 0.0f;
            }
        }
    }
}

/* ---------------------------------------------------------------- Menon -- */

typedef struct { int64_t x, y, w, h; } box;

/* Rectangle r grown by per-side halos and clipped to the active area,
 * MenonDemosaic.cpp:23-31. */
static box grow(box r, ro_rect a, int l, int t, int rt, int b) {
    int64_t x0 = r.x - l, y0 = r.y - t, x1 = r.x + r.w + rt, y1 = r.y + r.h + b;
    if (x0 < a.x) x0 = a.x;
    if (y0 < a.y) y0 = a.y;
    if (x1 > (int64_t)a.x + a.width) x1 = (int64_t)a.x + a.width;
    if (y1 > (int64_t)a.y + a.height) y1 = (int64_t)a.y + a.height;
    box o = {x0, y0, x1 - x0, y1 - y0};
    return o;
}

typedef struct { box b; double* v; } dplane;

static double* dat(dplane* p, int64_t x, int64_t y) {
    if (x < p->b.x || y < p->b.y || x >= p->b.x + p->b.w || y >= p->b.y + p->b.h) abort();
    return p->v + (size_t)(y - p->b.y) * (size_t)p->b.w + (size_t)(x - p->b.x);
}

/* Whole-sample reflection about the active-area edges, MenonDemosaic.cpp:60-66. */
static int64_t reflect(int64_t v, int64_t origin, int64_t size) {
    if (size == 1) return origin;
    const int64_t period = 2 * (size - 1);
    int64_t f = (v - origin) % period;
    if (f < 0) f += period;
    return origin + (f < size ? f : period - f);
}

/* Plane read with mirroring outside the active area, MenonDemosaic.cpp:67-72. */
static double mget(dplane* p, const ro_rect* a, int64_t x, int64_t y) {
    if (!inside(a, x, y)) { x = reflect(x, a->x, a->width); y = reflect(y, a->y, a->height); }
    return *dat(p, x, y);
}

static dplane mkplane(box b) {
    dplane p;
    p.b = b;
    p.v = (double*)calloc((size_t)(b.w > 0 ? b.w : 1) * (size_t)(b.h > 0 ? b.h : 1), sizeof(double));
    return p;
}

/* One output block, MenonDemosaic.cpp:89-174.  The raw plane is stored as
 * float values widened to double: the reference keeps it as float and widens
 * on every read, which is the same value. */
/* FP64 test entry inputs: raw values and frozen decisions, area-relative planes. */
typedef struct {
    const double* raw;      /* normalised samples (aw x ah) replacing the codes, or NULL */
    const uint8_t* hz_in;   /* classifier decisions to use (1 = horizontal), or NULL */
    uint8_t* hz_out;        /* decisions taken, or NULL */
    double* outd;           /* FP64 output (request-relative), or NULL for float out */
} menon_io;

static void menon_block(const ro_bayer* m, const uint16_t* s, box r, ro_rect out_rect, float* out,
                        const menon_io* io) {
    const ro_rect a = m->area;
    dplane raw = mkplane(grow(r, a, 6, 6, 6, 6));
    const box dirb = grow(r, a, 4, 4, 4, 4);
    dplane gh = mkplane(dirb), gv = mkplane(dirb), ch = mkplane(dirb), cv = mkplane(dirb);
    const box gb = grow(r, a, 2, 2, 2, 2);
    dplane green = mkplane(gb), horiz = mkplane(gb);
    dplane dh = mkplane(grow(gb, a, 2, 2, 0, 2)), dv = mkplane(grow(gb, a, 2, 2, 2, 0));
    const box cb = grow(r, a, 1, 1, 1, 1);
    dplane red = mkplane(cb), blue = mkplane(cb);
    int64_t x, y;

    for (y = raw.b.y; y < raw.b.y + raw.b.h; ++y)
        for (x = raw.b.x; x < raw.b.x + raw.b.w; ++x)
            *dat(&raw, x, y) = io && io->raw ? io->raw[(size_t)(y - a.y) * a.width + (size_t)(x - a.x)]
                                             : (double)normalized(m, s, x, y);

    for (y = dirb.y; y < dirb.y + dirb.h; ++y)
        for (x = dirb.x; x < dirb.x + dirb.w; ++x) {
            const double own = *dat(&raw, x, y);
            if (color_of(m, x, y) == 1) {
                *dat(&gh, x, y) = own; *dat(&gv, x, y) = own;
                continue;
            }
            const double h = 0.5 * (mget(&raw, &a, x - 1, y) + mget(&raw, &a, x + 1, y)) + 0.5 * own
                           - 0.25 * (mget(&raw, &a, x - 2, y) + mget(&raw, &a, x + 2, y));
            const double v = 0.5 * (mget(&raw, &a, x, y - 1) + mget(&raw, &a, x, y + 1)) + 0.5 * own
                           - 0.25 * (mget(&raw, &a, x, y - 2) + mget(&raw, &a, x, y + 2));
            *dat(&gh, x, y) = h; *dat(&gv, x, y) = v;
            *dat(&ch, x, y) = own - h; *dat(&cv, x, y) = own - v;
        }

    for (y = dh.b.y; y < dh.b.y + dh.b.h; ++y)
        for (x = dh.b.x; x < dh.b.x + dh.b.w; ++x)
            *dat(&dh, x, y) = fabs(*dat(&ch, x, y) - mget(&ch, &a, x + 2, y));
    for (y = dv.b.y; y < dv.b.y + dv.b.h; ++y)
        for (x = dv.b.x; x < dv.b.x + dv.b.w; ++x)
            *dat(&dv, x, y) = fabs(*dat(&cv, x, y) - mget(&cv, &a, x, y + 2));

    /* Classifier taps (dx, dy, weight), MenonDemosaic.cpp:131-132; the vertical
     * sum reads dv at the transposed offsets. */
    static const int tap[8][3] = {{0, 2, 1}, {-2, 2, 1}, {-1, 1, 1}, {0, 0, 3},
                                  {-2, 0, 3}, {-1, -1, 1}, {0, -2, 1}, {-2, -2, 1}};
    for (y = gb.y; y < gb.y + gb.h; ++y)
        for (x = gb.x; x < gb.x + gb.w; ++x) {
            if (color_of(m, x, y) == 1) {
                *dat(&horiz, x, y) = 1.0; *dat(&green, x, y) = *dat(&raw, x, y);
                continue;
            }
            double h = 0.0, v = 0.0;
            for (int t = 0; t < 8; ++t) {
                const double w = (double)tap[t][2];
                const int64_t hx = x + tap[t][0], hy = y + tap[t][1];
                const int64_t vx = x + tap[t][1], vy = y + tap[t][0];
                h += w * (inside(&a, hx, hy) ? *dat(&dh, hx, hy) : 0.0);
                v += w * (inside(&a, vx, vy) ? *dat(&dv, vx, vy) : 0.0);
            }
            int hz = v >= h;
            const size_t ai = (size_t)(y - a.y) * a.width + (size_t)(x - a.x);
            if (io && io->hz_in) hz = io->hz_in[ai] != 0;
            if (io && io->hz_out) io->hz_out[ai] = (uint8_t)hz;
            *dat(&horiz, x, y) = hz ? 1.0 : 0.0;
            *dat(&green, x, y) = hz ? *dat(&gh, x, y) : *dat(&gv, x, y);
        }

#define MEAN(P, X, Y, DX, DY) (0.5 * (mget((P), &a, (X) - (DX), (Y) - (DY)) + mget((P), &a, (X) + (DX), (Y) + (DY))))
    for (y = cb.y; y < cb.y + cb.h; ++y)
        for (x = cb.x; x < cb.x + cb.w; ++x) {
            const int own = color_of(m, x, y);
            if (own == 0) *dat(&red, x, y) = *dat(&raw, x, y);
            if (own == 2) *dat(&blue, x, y) = *dat(&raw, x, y);
            if (own == 1) {
                const int rh = color_of(m, x + 1, y) == 0;
                const int dx = rh ? 1 : 0, dy = rh ? 0 : 1;
                *dat(&red, x, y) = *dat(&green, x, y) + MEAN(&raw, x, y, dx, dy) - MEAN(&green, x, y, dx, dy);
                *dat(&blue, x, y) = *dat(&green, x, y) + MEAN(&raw, x, y, dy, dx) - MEAN(&green, x, y, dy, dx);
            }
        }

    for (y = r.y; y < r.y + r.h; ++y)
        for (x = r.x; x < r.x + r.w; ++x) {
            const int own = color_of(m, x, y);
            const int hz = *dat(&horiz, x, y) != 0.0;
            const int dx = hz ? 1 : 0, dy = hz ? 0 : 1;
            double rr = *dat(&red, x, y), bb = *dat(&blue, x, y);
            if (own == 2) rr = bb + MEAN(&red, x, y, dx, dy) - MEAN(&blue, x, y, dx, dy);
            if (own == 0) bb = rr + MEAN(&blue, x, y, dx, dy) - MEAN(&red, x, y, dx, dy);
            const size_t o = ((size_t)(y - out_rect.y) * out_rect.width + (size_t)(x - out_rect.x)) * 3;
            if (io && io->outd) {
                io->outd[o] = rr; io->outd[o + 1] = *dat(&green, x, y); io->outd[o + 2] = bb;
                continue;
            }
            out[o] = (float)rr;
            out[o + 1] = (float)*dat(&green, x, y);
            out[o + 2] = (float)bb;
        }
#undef MEAN
    free(raw.v); free(gh.v); free(gv.v); free(ch.v); free(cv.v);
    free(green.v); free(horiz.v); free(dh.v); free(dv.v); free(red.v); free(blue.v);
}

/* render_menon_base MenonDemosaic.cpp:177-218: 256x256 output blocks. */
static void menon_blocks(const ro_bayer* m, const uint16_t* s, ro_rect r, ro_rect out_rect, float* out,
                         const menon_io* io) {
    const uint32_t B = 256;
    for (uint64_t by = r.y; by < (uint64_t)r.y + r.height; by += B)
        for (uint64_t bx = r.x; bx < (uint64_t)r.x + r.width; bx += B) {
            box b;
            b.x = (int64_t)bx; b.y = (int64_t)by;
            b.w = (int64_t)(((uint64_t)r.x + r.width - bx) < B ? ((uint64_t)r.x + r.width - bx) : B);
            b.h = (int64_t)(((uint64_t)r.y + r.height - by) < B ? ((uint64_t)r.y + r.height - by) : B);
            menon_block(m, s, b, out_rect, out, io);
        }
}

int ro_demosaic_menon(const ro_bayer* m, const uint16_t* s, ro_rect r, float* out) {
    if (m->area.width <= 1 || m->area.height <= 1) return -1;
    if (!r.width || !r.height) return 0;
    menon_blocks(m, s, r, r, out, NULL);
    return 0;
}

int ro_menon_f64(const ro_bayer* m, const double* raw, ro_rect r, const uint8_t* hz_in, uint8_t* hz_out,
                 double* out) {
    if (m->area.width <= 1 || m->area.height <= 1) return -1;
    if (!r.width || !r.height) return 0;
    menon_io io = {raw, hz_in, hz_out, out};
    menon_blocks(m, NULL, r, r, NULL, &io);
    return 0;
}

int ro_demosaic_menon_bands(const ro_bayer* m, const uint16_t* s, ro_rect r, float* out, int threads) {
    if (m->area.width <= 1 || m->area.height <= 1) return -1;
    if (!r.width || !r.height) return 0;
    if (threads < 1) threads = 1;
    const uint32_t band = (r.height + (uint32_t)threads - 1) / (uint32_t)threads;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static, 1)
#endif
    for (int t = 0; t < threads; ++t) {
        const uint32_t y0 = (uint32_t)t * band;
        if (y0 >= r.height) continue;
        const uint32_t h = r.height - y0 < band ? r.height - y0 : band;
        ro_rect sub = {r.x, r.y + y0, r.width, h};
        menon_blocks(m, s, sub, r, out, NULL);
    }
    return 0;
}

/* ----------------------------------------------------------- point ops -- */

void ro_scale3(float* rgb, size_t pixels, const float g[3]) {
    for (size_t i = 0; i < pixels; ++i) {
        rgb[3 * i] *= g[0]; rgb[3 * i + 1] *= g[1]; rgb[3 * i + 2] *= g[2];
    }
}

void ro_scale(float* v, size_t n, float k) {
    for (size_t i = 0; i < n; ++i) v[i] *= k;
}

void ro_mat3(float* rgb, size_t pixels, const float mt[9]) {
    for (size_t i = 0; i < pixels; ++i) {
        const float r = rgb[3 * i], g = rgb[3 * i + 1], b = rgb[3 * i + 2];
        rgb[3 * i] = mt[0] * r + mt[1] * g + mt[2] * b;
        rgb[3 * i + 1] = mt[3] * r + mt[4] * g + mt[5] * b;
        rgb[3 * i + 2] = mt[6] * r + mt[7] * g + mt[8] * b;
    }
}

void ro_reduce_mip(const float* nat, uint32_t nw, uint32_t nh, uint32_t mip,
                   uint32_t ow, uint32_t oh, float* out) {
    const uint32_t s = 1u << mip;
    for (uint32_t y = 0; y < oh; ++y)
        for (uint32_t x = 0; x < ow; ++x) {
            double sum[3] = {0, 0, 0};
            uint32_t count = 0;
            const uint64_t ye = (uint64_t)(y + 1) * s < nh ? (uint64_t)(y + 1) * s : nh;
            const uint64_t xe = (uint64_t)(x + 1) * s < nw ? (uint64_t)(x + 1) * s : nw;
            for (uint64_t sy = (uint64_t)y * s; sy < ye; ++sy)
                for (uint64_t sx = (uint64_t)x * s; sx < xe; ++sx) {
                    const float* p = nat + ((size_t)sy * nw + sx) * 3;
                    sum[0] += p[0]; sum[1] += p[1]; sum[2] += p[2];
                    ++count;
                }
            float* o = out + ((size_t)y * ow + x) * 3;
            for (int c = 0; c < 3; ++c) o[c] = (float)(sum[c] / count);
        }
}

void ro_tone_curve(float* v, size_t n, float shoulder, float inverse_gamma) {
    for (size_t i = 0; i < n; ++i) {
        const float linear = v[i];
        const float mag = fabsf(linear);
        const float compressed = mag / (mag + shoulder);
        const float mapped = inverse_gamma == 1.0f ? compressed : powf(compressed, inverse_gamma);
        v[i] = copysignf(mapped, linear);
    }
}

int ro_clip(float* v, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        if (!isfinite(v[i])) return -1;
        v[i] = v[i] < 0.0f ? 0.0f : (v[i] > 1.0f ? 1.0f : v[i]);
    }
    return 0;
}

int ro_srgb_encode(float* v, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (!isfinite(v[i])) return -1;
    for (size_t i = 0; i < n; ++i) {
        const float l = v[i] < 0.0f ? 0.0f : (v[i] > 1.0f ? 1.0f : v[i]);
        v[i] = l <= 0.0031308f ? 12.92f * l : 1.055f * powf(l, 1.0f / 2.4f) - 0.055f;
    }
    return 0;
}

void ro_srgb_u8(const float* e, size_t n, uint8_t* out) {
    for (size_t i = 0; i < n; ++i) {
        const float c = e[i] < 0.0f ? 0.0f : (e[i] > 1.0f ? 1.0f : e[i]);
        out[i] = (uint8_t)(c * 255.0f + 0.5f);
    }
}

float ro_exp2f(float stops) { return exp2f(stops); }

/* --------------------------------------------------------------- curves -- */

typedef struct {
    const double *x, *y;
    uint32_t n;
    int identity;
    double slope[256], ctl0[256], ctl1[256];
} curve;

/* ExtendedCurvesNode constructor, ToneOps.cpp:795-816. */
static void curve_prepare(curve* c, const double* xs, const double* ys, uint32_t n, int cubic) {
    c->x = xs; c->y = ys; c->n = n; c->identity = 1;
    for (uint32_t i = 0; i < n; ++i) if (xs[i] != ys[i]) c->identity = 0;
    for (uint32_t i = 1; i < n; ++i) c->slope[i - 1] = (ys[i] - ys[i - 1]) / (xs[i] - xs[i - 1]);
    if (!cubic) return;
    double tan[257];
    tan[0] = c->slope[0];
    for (uint32_t i = 1; i + 1 < n; ++i) {
        const double a = c->slope[i - 1], b = c->slope[i];
        tan[i] = (a == 0 || b == 0 || signbit(a) != signbit(b)) ? 0.0 : copysign(fmin(fabs(a), fabs(b)), a);
    }
    tan[n - 1] = c->slope[n - 2];
    for (uint32_t i = 0; i + 1 < n; ++i) {
        const double h = xs[i + 1] - xs[i];
        c->ctl0[i] = ys[i] + (h * tan[i]) / 3;
        c->ctl1[i] = ys[i + 1] - (h * tan[i + 1]) / 3;
    }
}

/* curve_cubic_blend ToneOps.cpp:788-792. */
static double blend(double a, double b, double t) {
    if (a == b || t == 0) return a;
    if (t == 1) return b;
    return t <= .5 ? a + t * (b - a) : b + (1 - t) * (a - b);
}

/* ExtendedCurvesNode::map ToneOps.cpp:820-831. */
static double curve_map(const curve* c, int cubic, double v) {
    if (c->identity) return v;
    uint32_t next = 0;
    while (next < c->n && !(v < c->x[next])) ++next;   /* upper_bound */
    const uint32_t left = next == 0 ? 0 : next - 1;
    const uint32_t i = left < c->n - 2 ? left : c->n - 2;
    if (v == c->x[left]) return c->y[left];
    if (!cubic || v < c->x[0] || v > c->x[c->n - 1])
        return c->slope[i] == 0 ? c->y[left] : c->y[left] + (v - c->x[left]) * c->slope[i];
    const double t = (v - c->x[i]) / (c->x[i + 1] - c->x[i]);
    const double a = blend(c->y[i], c->ctl0[i], t), b = blend(c->ctl0[i], c->ctl1[i], t),
                 d = blend(c->ctl1[i], c->y[i + 1], t);
    return blend(blend(a, b, t), blend(b, d, t), t);
}

int ro_curves(float* rgb, size_t pixels, const double* xs, const double* ys,
              const uint32_t off[5], int cubic) {
    curve* c = (curve*)malloc(sizeof(curve) * 4);
    for (int k = 0; k < 4; ++k) curve_prepare(&c[k], xs + off[k], ys + off[k], off[k + 1] - off[k], cubic);
    int status = 0;
    for (size_t i = 0; i < pixels * 3 && !status; ++i)
        if (!isfinite(rgb[i])) status = -1;
    for (size_t i = 0; i < pixels * 3 && !status; ++i) {
        const int ch = (int)(i % 3) + 1;
        if (c[0].identity && c[ch].identity) continue;
        const double master = curve_map(&c[0], cubic, rgb[i]);
        if (!isfinite(master)) { status = -1; break; }
        const double value = curve_map(&c[ch], cubic, master);
        if (!isfinite(value) || fabs(value) > FLT_MAX) { status = -1; break; }
        rgb[i] = (float)value;
    }
    free(c);
    return status;
}

/* -------------------------------------------------------- guided filter -- */

static ro_rect clip_grow(ro_rect r, ro_rect img, uint32_t rad) {
    int64_t x0 = (int64_t)r.x - rad, y0 = (int64_t)r.y - rad;
    int64_t x1 = (int64_t)r.x + r.width + rad, y1 = (int64_t)r.y + r.height + rad;
    if (x0 < img.x) x0 = img.x;
    if (y0 < img.y) y0 = img.y;
    if (x1 > (int64_t)img.x + img.width) x1 = (int64_t)img.x + img.width;
    if (y1 > (int64_t)img.y + img.height) y1 = (int64_t)img.y + img.height;
    ro_rect o = {(uint32_t)x0, (uint32_t)y0, (uint32_t)(x1 - x0), (uint32_t)(y1 - y0)};
    return o;
}

typedef struct { double my, mp[3], slope[3]; } gf_coef;

/* working_y_guided_filter_rgb SpatialOps.cpp:168-287: ordered, centred FP64. */
static int gf_rows(const float* in, ro_rect needed, ro_rect output, ro_rect image,
                   uint32_t rad, double eps, double wr, double wb, float* out, ro_rect out_full) {
    if (rad == 0) {
        for (uint32_t y = 0; y < output.height; ++y)
            for (uint32_t x = 0; x < output.width; ++x)
                for (int c = 0; c < 3; ++c)
                    out[((size_t)(output.y - out_full.y + y) * out_full.width + (output.x - out_full.x + x)) * 3 + c] =
                        in[((size_t)(output.y - needed.y + y) * needed.width + (output.x - needed.x + x)) * 3 + c];
        return 0;
    }
    const ro_rect centers = clip_grow(output, image, rad);
    gf_coef* table = (gf_coef*)malloc(sizeof(gf_coef) * (size_t)centers.width * centers.height);
    int status = 0;
#define PIX(X, Y) (in + ((size_t)((Y) - needed.y) * needed.width + (size_t)((X) - needed.x)) * 3)
#define YDELTA(D) (((D)[1] + wr * ((D)[0] - (D)[1])) + wb * ((D)[2] - (D)[1]))
    for (uint64_t ky = centers.y; ky < (uint64_t)centers.y + centers.height && !status; ++ky)
        for (uint64_t kx = centers.x; kx < (uint64_t)centers.x + centers.width; ++kx) {
            ro_rect one = {(uint32_t)kx, (uint32_t)ky, 1, 1};
            const ro_rect w = clip_grow(one, image, rad);
            const double count = (double)((uint64_t)w.width * w.height);
            const float* pk = PIX(kx, ky);
            double sg = 0, sp[3] = {0, 0, 0};
            for (uint64_t y = w.y; y < (uint64_t)w.y + w.height; ++y)
                for (uint64_t x = w.x; x < (uint64_t)w.x + w.width; ++x) {
                    const float* pj = PIX(x, y);
                    double d[3];
                    for (int c = 0; c < 3; ++c) d[c] = (double)pj[c] - (double)pk[c];
                    sg += YDELTA(d);
                    for (int c = 0; c < 3; ++c) sp[c] += d[c];
                }
            gf_coef co;
            co.my = sg / count;
            for (int c = 0; c < 3; ++c) co.mp[c] = sp[c] / count;
            if (!isfinite(co.my) || !isfinite(co.mp[0]) || !isfinite(co.mp[1]) || !isfinite(co.mp[2])) { status = -1; break; }
            double vv = 0, cov[3] = {0, 0, 0};
            for (uint64_t y = w.y; y < (uint64_t)w.y + w.height; ++y)
                for (uint64_t x = w.x; x < (uint64_t)w.x + w.width; ++x) {
                    const float* pj = PIX(x, y);
                    double d[3];
                    for (int c = 0; c < 3; ++c) d[c] = (double)pj[c] - (double)pk[c];
                    const double dy = YDELTA(d) - co.my;
                    vv += dy * dy;
                    for (int c = 0; c < 3; ++c) {
                        const double dp = d[c] - co.mp[c];
                        cov[c] += dy * dp;
                    }
                }
            const double variance = vv / count, denominator = variance + eps;
            if (!isfinite(variance) || !isfinite(denominator) || variance < 0 || denominator < eps) { status = -1; break; }
            for (int c = 0; c < 3; ++c) {
                co.slope[c] = (cov[c] / count) / denominator;
                if (!isfinite(co.slope[c])) status = -1;
            }
            table[(size_t)(ky - centers.y) * centers.width + (size_t)(kx - centers.x)] = co;
        }
    for (uint64_t y = output.y; y < (uint64_t)output.y + output.height && !status; ++y)
        for (uint64_t x = output.x; x < (uint64_t)output.x + output.width; ++x) {
            ro_rect one = {(uint32_t)x, (uint32_t)y, 1, 1};
            const ro_rect w = clip_grow(one, image, rad);
            const double count = (double)((uint64_t)w.width * w.height);
            const float* pi = PIX(x, y);
            double corr[3] = {0, 0, 0};
            for (uint64_t ky = w.y; ky < (uint64_t)w.y + w.height; ++ky)
                for (uint64_t kx = w.x; kx < (uint64_t)w.x + w.width; ++kx) {
                    const gf_coef* co = &table[(size_t)(ky - centers.y) * centers.width + (size_t)(kx - centers.x)];
                    const float* pk = PIX(kx, ky);
                    double d[3], rev[3];
                    for (int c = 0; c < 3; ++c) { d[c] = (double)pk[c] - (double)pi[c]; rev[c] = (double)pi[c] - (double)pk[c]; }
                    const double guide = YDELTA(rev) - co->my;
                    for (int c = 0; c < 3; ++c) {
                        double term = d[c] + co->mp[c];
                        if (co->slope[c] != 0) term += co->slope[c] * guide;
                        corr[c] += term;
                    }
                }
            float* po = out + ((size_t)(y - out_full.y) * out_full.width + (size_t)(x - out_full.x)) * 3;
            for (int c = 0; c < 3; ++c) {
                const double delta = corr[c] / count;
                if (!isfinite(delta)) { status = -1; break; }
                if (delta == 0) { po[c] = pi[c]; continue; }
                const double value = (double)pi[c] + delta;
                if (!isfinite(value) || value < -(double)FLT_MAX || value > (double)FLT_MAX) { status = -1; break; }
                po[c] = (float)value;
            }
        }
#undef PIX
#undef YDELTA
    free(table);
    return status;
}

/* FP64 I/O variant for finite differences: same ordered algorithm, the input
 * rounded to float32 is replaced by double values and the output is not
 * rounded (SpatialOps.cpp:168-287 arithmetic, without the float32 casts). */
int ro_guided_filter_f64(const double* in, ro_rect needed, ro_rect output, ro_rect image,
                         uint32_t rad, double eps, double wr, double wb, double* out) {
    const ro_rect centers = clip_grow(output, image, rad);
    gf_coef* table = (gf_coef*)malloc(sizeof(gf_coef) * (size_t)centers.width * centers.height);
#define PIXD(X, Y) (in + ((size_t)((Y) - needed.y) * needed.width + (size_t)((X) - needed.x)) * 3)
#define YD(D) (((D)[1] + wr * ((D)[0] - (D)[1])) + wb * ((D)[2] - (D)[1]))
    for (uint64_t ky = centers.y; ky < (uint64_t)centers.y + centers.height; ++ky)
        for (uint64_t kx = centers.x; kx < (uint64_t)centers.x + centers.width; ++kx) {
            ro_rect one = {(uint32_t)kx, (uint32_t)ky, 1, 1};
            const ro_rect w = clip_grow(one, image, rad);
            const double count = (double)((uint64_t)w.width * w.height);
            const double* pk = PIXD(kx, ky);
            double sg = 0, sp[3] = {0, 0, 0};
            for (uint64_t y = w.y; y < (uint64_t)w.y + w.height; ++y)
                for (uint64_t x = w.x; x < (uint64_t)w.x + w.width; ++x) {
                    const double* pj = PIXD(x, y);
                    double d[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
                    sg += YD(d);
                    for (int c = 0; c < 3; ++c) sp[c] += d[c];
                }
            gf_coef co;
            co.my = sg / count;
            for (int c = 0; c < 3; ++c) co.mp[c] = sp[c] / count;
            double vv = 0, cov[3] = {0, 0, 0};
            for (uint64_t y = w.y; y < (uint64_t)w.y + w.height; ++y)
                for (uint64_t x = w.x; x < (uint64_t)w.x + w.width; ++x) {
                    const double* pj = PIXD(x, y);
                    double d[3] = {pj[0] - pk[0], pj[1] - pk[1], pj[2] - pk[2]};
                    const double dy = YD(d) - co.my;
                    vv += dy * dy;
                    for (int c = 0; c < 3; ++c) cov[c] += dy * (d[c] - co.mp[c]);
                }
            const double denominator = vv / count + eps;
            for (int c = 0; c < 3; ++c) co.slope[c] = (cov[c] / count) / denominator;
            table[(size_t)(ky - centers.y) * centers.width + (size_t)(kx - centers.x)] = co;
        }
    for (uint64_t y = output.y; y < (uint64_t)output.y + output.height; ++y)
        for (uint64_t x = output.x; x < (uint64_t)output.x + output.width; ++x) {
            ro_rect one = {(uint32_t)x, (uint32_t)y, 1, 1};
            const ro_rect w = clip_grow(one, image, rad);
            const double count = (double)((uint64_t)w.width * w.height);
            const double* pi = PIXD(x, y);
            double corr[3] = {0, 0, 0};
            for (uint64_t ky = w.y; ky < (uint64_t)w.y + w.height; ++ky)
                for (uint64_t kx = w.x; kx < (uint64_t)w.x + w.width; ++kx) {
                    const gf_coef* co = &table[(size_t)(ky - centers.y) * centers.width + (size_t)(kx - centers.x)];
                    const double* pk = PIXD(kx, ky);
                    double rev[3] = {pi[0] - pk[0], pi[1] - pk[1], pi[2] - pk[2]};
                    const double guide = YD(rev) - co->my;
                    for (int c = 0; c < 3; ++c) corr[c] += (pk[c] - pi[c]) + co->mp[c] + co->slope[c] * guide;
                }
            double* po = out + ((size_t)(y - output.y) * output.width + (size_t)(x - output.x)) * 3;
            for (int c = 0; c < 3; ++c) po[c] = pi[c] + corr[c] / count;
        }
#undef PIXD
#undef YD
    free(table);
    return 0;
}

/* Master curve map in FP64 (ExtendedCurvesNode::map ToneOps.cpp:820-831). */
void ro_curve_map_f64(const double* xs, const double* ys, uint32_t n, int cubic, const double* in, double* out,
                      size_t count) {
    curve* c = (curve*)malloc(sizeof(curve));
    curve_prepare(c, xs, ys, n, cubic);
    c->identity = 0;
    for (size_t i = 0; i < count; ++i) out[i] = curve_map(c, cubic, in[i]);
    free(c);
}

int ro_guided_filter(const float* in, ro_rect needed, ro_rect output, ro_rect image,
                     uint32_t rad, double eps, double wr, double wb, float* out) {
    for (size_t i = 0; i < (size_t)needed.width * needed.height * 3; ++i)
        if (!isfinite(in[i])) return -1;
    return gf_rows(in, needed, output, image, rad, eps, wr, wb, out, output);
}

int ro_guided_filter_bands(const float* in, ro_rect needed, ro_rect output, ro_rect image,
                           uint32_t rad, double eps, double wr, double wb, float* out, int threads) {
    if (threads < 1) threads = 1;
    const uint32_t band = (output.height + (uint32_t)threads - 1) / (uint32_t)threads;
    int status = 0;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static, 1) reduction(|:status)
#endif
    for (int t = 0; t < threads; ++t) {
        const uint32_t y0 = (uint32_t)t * band;
        if (y0 >= output.height) continue;
        ro_rect sub = {output.x, output.y + y0, output.width,
                       output.height - y0 < band ? output.height - y0 : band};
        status |= gf_rows(in, needed, sub, image, rad, eps, wr, wb, out, output) ? 1 : 0;
    }
    return status ? -1 : 0;
}

/* ----------------------------------------------------- exact division -- */

static float nearest_recip(uint32_t d) {
    const float y = (float)(1.0 / (double)d);
    float best = y;
    double err = fabs((double)y * d - 1.0);
    const float c[2] = {nextafterf(y, 0.0f), nextafterf(y, 2.0f)};
    for (int i = 0; i < 2; ++i) {
        const double e = fabs((double)c[i] * d - 1.0);
        if (e < err) { err = e; best = c[i]; }
    }
    return best;
}

#define DIV_BODY                                                            \
    long long bad = 0;                                                      \
    _Pragma("omp parallel for schedule(dynamic, 64) reduction(+ : bad)")    \
    for (long long d = d_lo; d <= (long long)d_hi; ++d) {                   \
        const float fd = (float)d, y = nearest_recip((uint32_t)d);          \
        for (uint32_t n = 0; n <= 65535u; ++n) {                            \
            const float fn = (float)n;                                      \
            volatile float q = fn / fd;                                     \
            const float q0 = fn * y;                                        \
            const float q1 = planted ? q0 : fmaf(fmaf(-q0, fd, fn), y, q0); \
            bad += q1 != q;                                                 \
        }                                                                   \
    }                                                                       \
    return bad;

__attribute__((target("fma"))) static long long div_mismatches_fma(uint32_t d_lo, uint32_t d_hi, int planted) {
    DIV_BODY
}
static long long div_mismatches_libm(uint32_t d_lo, uint32_t d_hi, int planted) { DIV_BODY }

long long ro_division_mismatches(uint32_t d_lo, uint32_t d_hi, int planted) {
    if (d_lo < 1) d_lo = 1;
    if (d_hi > 65535) d_hi = 65535;
    if (__builtin_cpu_supports("fma")) return div_mismatches_fma(d_lo, d_hi, planted);
    return div_mismatches_libm(d_lo, d_hi, planted);
}


/* ---- local statistics (SpatialOps.cpp:48-121) ---- */

typedef struct { uint32_t n; double mean, m2; } ro_moment;

static void moment_add(ro_moment* m, float v) {   /* SpatialOps.cpp:48-55 */
    if (!isfinite(v)) return;
    ++m->n;
    const double d = (double)v - m->mean;
    m->mean += d / m->n;
    m->m2 += d * ((double)v - m->mean);
}

static void moment_merge(ro_moment* a, const ro_moment* b) {   /* SpatialOps.cpp:56-64 */
    if (!b->n) return;
    if (!a->n) { *a = *b; return; }
    const uint32_t total = a->n + b->n;
    const double d = b->mean - a->mean;
    a->m2 += b->m2 + d * d * ((double)a->n * b->n / total);
    a->mean += d * ((double)b->n / total);
    a->n = total;
}

void ro_local_statistics(const float* in, ro_rect needed, ro_rect output, ro_rect image, uint32_t radius,
                         double* mean, double* variance, uint32_t* count) {
    const size_t hw = output.width;
    ro_moment* h = (ro_moment*)calloc((size_t)needed.height * hw * 3, sizeof(ro_moment));
    for (uint32_t row = 0; row < needed.height; ++row)
        for (uint32_t col = 0; col < output.width; ++col) {
            const int64_t x = (int64_t)output.x + col;
            const int64_t left = x - radius > (int64_t)image.x ? x - radius : (int64_t)image.x;
            const int64_t lastx = (int64_t)image.x + image.width - 1;
            const int64_t right = x + radius < lastx ? x + radius : lastx;
            for (int64_t xx = left; xx <= right; ++xx) {
                const float* p = in + ((size_t)row * needed.width + (size_t)(xx - needed.x)) * 3;
                for (int c = 0; c < 3; ++c) moment_add(&h[((size_t)row * hw + col) * 3 + c], p[c]);
            }
        }
    for (uint32_t row = 0; row < output.height; ++row) {
        const int64_t y = (int64_t)output.y + row;
        const int64_t top = y - radius > (int64_t)image.y ? y - radius : (int64_t)image.y;
        const int64_t lasty = (int64_t)image.y + image.height - 1;
        const int64_t bottom = y + radius < lasty ? y + radius : lasty;
        for (uint32_t col = 0; col < output.width; ++col)
            for (int c = 0; c < 3; ++c) {
                ro_moment t = {0, 0.0, 0.0};
                for (int64_t yy = top; yy <= bottom; ++yy)
                    moment_merge(&t, &h[((size_t)(yy - needed.y) * hw + col) * 3 + c]);
                const size_t o = ((size_t)row * output.width + col) * 3 + c;
                count[o] = t.n;
                if (t.n) {
                    const double v = t.m2 / t.n;
                    mean[o] = t.mean;
                    variance[o] = 0.0 < v ? v : 0.0;
                } else {
                    mean[o] = NAN;
                    variance[o] = NAN;
                }
            }
    }
    free(h);
}
