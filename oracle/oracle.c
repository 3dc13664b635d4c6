/*
 * oracle.c -- CPU restatement used only as the checker (tests, smoke, the
 * cpu_baseline leg of bench.py).  Build: oracle/Makefile, -ffp-contract=off.
 * See oracle.h for the list of reference functions restated here.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ======================================================================
 * FP32 sin/cos: Cody-Waite three-part reduction by pi/2 followed by the
 * Taylor polynomials of sin and cos on [-pi/4, pi/4] (degree 9 and 10),
 * evaluated in Horner form.  The three parts of pi/2 are its leading 12
 * bits, the next 12 bits and the FP32-rounded remainder, so k*C1 and
 * k*C2 are exact for |k| < 2^12.
 * ====================================================================== */
static const float ORC_INV_PIO2 = 0.636619772367581343f; /* 2/pi */
static const float ORC_PIO2_1 = 1.5703125f;                /* 0x1.92p+0 */
static const float ORC_PIO2_2 = 4.837512969970703125e-4f;  /* 0x1.fb4p-12 */
static const float ORC_PIO2_3 = 7.549790126404332e-08f;    /* 0x1.4442d2p-24 */

static const float ORC_S1 = -1.66666666666666667e-1f; /* -1/3! */
static const float ORC_S2 = 8.33333333333333333e-3f;  /*  1/5! */
static const float ORC_S3 = -1.98412698412698413e-4f; /* -1/7! */
static const float ORC_S4 = 2.75573192239858907e-6f;  /*  1/9! */
static const float ORC_C1 = 4.16666666666666667e-2f;  /*  1/4! */
static const float ORC_C2 = -1.38888888888888889e-3f; /* -1/6! */
static const float ORC_C3 = 2.48015873015873016e-5f;  /*  1/8! */
static const float ORC_C4 = -2.75573192239858907e-7f; /* -1/10! */

void orc_sincosf(float x, float *s, float *c)
{
    float kf = rintf(x * ORC_INV_PIO2);
    int k = (int)kf;
    float r = x - kf * ORC_PIO2_1;
    r = r - kf * ORC_PIO2_2;
    r = r - kf * ORC_PIO2_3;
    float z = r * r;
    float ps = r + (r * z) * (ORC_S1 + z * (ORC_S2 + z * (ORC_S3 + z * ORC_S4)));
    float pc = 1.0f + z * (-0.5f + z * (ORC_C1 + z * (ORC_C2 + z * (ORC_C3 + z * ORC_C4))));
    switch (k & 3) {
    case 0: *s = ps; *c = pc; break;
    case 1: *s = pc; *c = -ps; break;
    case 2: *s = -ps; *c = -pc; break;
    default: *s = -pc; *c = ps; break;
    }
}

/* acos on [0,1]: Abramowitz & Stegun 4.4.46 (Hastings), |err| <= 2e-8:
 * acos(x) = sqrt(1-x) * (a0 + a1 x + ... + a7 x^7). */
float orc_acosf(float x)
{
    float p = -0.0012624911f;
    p = p * x + 0.0066700901f;
    p = p * x + -0.0170881256f;
    p = p * x + 0.0308918810f;
    p = p * x + -0.0501743046f;
    p = p * x + 0.0889789874f;
    p = p * x + -0.2145988016f;
    p = p * x + 1.5707963050f;
    return sqrtf(1.0f - x) * p;
}

/* ======================================================================
 * Arm specification (7-DOF serial arm of the benchmark workloads):
 *  - base at the world origin, joint 0 at the origin;
 *  - joint i rotates about local z for even i and local y for odd i;
 *  - R_0 = I * Rot(q_0), R_i = R_{i-1} * Rot_axis(i)(q_i);
 *  - link i runs from p_i along the z column of R_i for L = 0.3 m;
 *    p_{i+1} = p_i + L * Z_i;
 *  - link i carries 4 spheres of radius 0.05 m at p_i + (L*f_j) * Z_i,
 *    f = {1/8, 3/8, 5/8, 7/8} (offsets 0.0375, 0.1125, 0.1875, 0.2625);
 *  - rotation update, with X,Y,Z the columns of R and (s,c) = sincos(q):
 *      z joint: X' = X*c + Y*s,  Y' = Y*c - X*s
 *      y joint: X' = X*c - Z*s,  Z' = X*s + Z*c
 *    evaluated per component in exactly that order;
 *  - centres are written sphere-major, link by link.
 * ====================================================================== */
static const float ORC_LINK = 0.3f;
static const float ORC_OFF[4] = {0.0375f, 0.1125f, 0.1875f, 0.2625f};
static const float ORC_ROBOT_R = 0.05f;

void orc_arm_spheres(const float q[ORC_ARM_DOF], float centres[ORC_ARM_SPHERES][3])
{
    float X[3] = {1.0f, 0.0f, 0.0f}, Y[3] = {0.0f, 1.0f, 0.0f}, Z[3] = {0.0f, 0.0f, 1.0f};
    float p[3] = {0.0f, 0.0f, 0.0f};
    for (int i = 0; i < ORC_ARM_DOF; ++i) {
        float s, c;
        orc_sincosf(q[i], &s, &c);
        if ((i & 1) == 0) {
            for (int k = 0; k < 3; ++k) {
                float nx = X[k] * c + Y[k] * s;
                float ny = Y[k] * c - X[k] * s;
                X[k] = nx;
                Y[k] = ny;
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                float nx = X[k] * c - Z[k] * s;
                float nz = X[k] * s + Z[k] * c;
                X[k] = nx;
                Z[k] = nz;
            }
        }
        for (int j = 0; j < 4; ++j)
            for (int k = 0; k < 3; ++k)
                centres[i * 4 + j][k] = p[k] + Z[k] * ORC_OFF[j];
        for (int k = 0; k < 3; ++k)
            p[k] = p[k] + Z[k] * ORC_LINK;
    }
}

/* Sphere/sphere: d = o - c per axis, d2 = (dx*dx + dy*dy) + dz*dz,
 * collision iff d2 < (R + r)*(R + r)  (strict, tests/resources/circles2D.h).
 * Sphere/box: lo = b - h, hi = b + h, p = min(max(c, lo), hi), d = c - p,
 * collision iff d2 < r*r. */
static int sphere_hits_sphere(const float c[3], const float *o)
{
    float rs = o[3] + ORC_ROBOT_R;
    float rr = rs * rs;
    float dx = o[0] - c[0], dy = o[1] - c[1], dz = o[2] - c[2];
    float d2 = (dx * dx + dy * dy) + dz * dz;
    return d2 < rr;
}

static int sphere_hits_box(const float c[3], const float *b)
{
    float d2 = 0.0f;
    float acc[3];
    for (int k = 0; k < 3; ++k) {
        float lo = b[k] - b[3 + k];
        float hi = b[k] + b[3 + k];
        float pk = fminf(fmaxf(c[k], lo), hi);
        float dk = c[k] - pk;
        acc[k] = dk * dk;
    }
    d2 = (acc[0] + acc[1]) + acc[2];
    return d2 < ORC_ROBOT_R * ORC_ROBOT_R;
}

int orc_state_valid(const orc_env *env, const float q[ORC_ARM_DOF], long *pairs)
{
    float cs[ORC_ARM_SPHERES][3];
    orc_arm_spheres(q, cs);
    long np = 0;
    for (int s = 0; s < ORC_ARM_SPHERES; ++s) {
        for (int o = 0; o < env->n_spheres; ++o) {
            ++np;
            if (sphere_hits_sphere(cs[s], env->spheres + 4 * o)) {
                if (pairs) *pairs = np;
                return 0;
            }
        }
        for (int o = 0; o < env->n_boxes; ++o) {
            ++np;
            if (sphere_hits_box(cs[s], env->boxes + 6 * o)) {
                if (pairs) *pairs = np;
                return 0;
            }
        }
    }
    if (pairs) *pairs = np;
    return 1;
}

int orc_state_valid_d(const orc_env *env, const double q[ORC_ARM_DOF], long *pairs)
{
    float qf[ORC_ARM_DOF];
    for (int i = 0; i < ORC_ARM_DOF; ++i)
        qf[i] = (float)q[i];
    return orc_state_valid(env, qf, pairs);
}

void orc_validate_states(const orc_env *env, const float *q_soa, long n, uint8_t *out)
{
    for (long i = 0; i < n; ++i) {
        float q[ORC_ARM_DOF];
        for (int j = 0; j < ORC_ARM_DOF; ++j)
            q[j] = q_soa[(long)j * n + i];
        out[i] = (uint8_t)orc_state_valid(env, q, NULL);
    }
}

/* ======================================================================
 * RealVectorStateSpace / StateSpace pieces (FP64, reference order)
 * ====================================================================== */
double orc_max_extent(const double *lo, const double *hi, int dim)
{
    double e = 0.0;
    for (int i = 0; i < dim; ++i) {
        double d = hi[i] - lo[i];
        e += d * d;
    }
    return sqrt(e);
}

double orc_longest_valid_segment(double max_extent, double fraction)
{
    return max_extent * fraction;
}

double orc_distance_d(const double *a, const double *b, int dim)
{
    double dist = 0.0;
    for (int i = 0; i < dim; ++i) {
        double diff = a[i] - b[i];
        dist += diff * diff;
    }
    return sqrt(dist);
}

void orc_interpolate_d(const double *a, const double *b, double t, int dim, double *out)
{
    for (int i = 0; i < dim; ++i)
        out[i] = a[i] + (b[i] - a[i]) * t;
}

unsigned orc_valid_segment_count(const double *a, const double *b, int dim, double lvs, unsigned factor)
{
    return factor * (unsigned)ceil(orc_distance_d(a, b, dim) / lvs);
}

/* DiscreteMotionValidator.cpp:106-134: a FIFO of [first, second] intervals
 * starting with (1, nd-1); pop, check mid=(first+second)/2, push the left
 * part then the right part. */
int orc_bfs_order(int nd, int *out)
{
    if (nd < 2) return 0;
    int cap = nd; /* at most nd-1 intervals ever live */
    int *qa = (int *)malloc(sizeof(int) * (size_t)cap * 2);
    int *qb = qa + cap;
    int head = 0, tail = 0, n = 0, live = 0;
    qa[tail] = 1; qb[tail] = nd - 1; tail = (tail + 1) % cap; live = 1;
    while (live > 0) {
        int a = qa[head], b = qb[head];
        int mid = (a + b) / 2;
        out[n++] = mid;
        head = (head + 1) % cap; --live;
        if (a < mid) { qa[tail] = a; qb[tail] = mid - 1; tail = (tail + 1) % cap; ++live; }
        if (b > mid) { qa[tail] = mid + 1; qb[tail] = b; tail = (tail + 1) % cap; ++live; }
    }
    free(qa);
    return n;
}

int orc_check_motion(const orc_env *env, const double a[ORC_ARM_DOF], const double b[ORC_ARM_DOF],
                     double lvs, unsigned factor, int *checked)
{
    int cnt = 1;
    if (!orc_state_valid_d(env, b, NULL)) {
        if (checked) *checked = cnt;
        return 0;
    }
    int nd = (int)orc_valid_segment_count(a, b, ORC_ARM_DOF, lvs, factor);
    int result = 1;
    if (nd >= 2) {
        int *order = (int *)malloc(sizeof(int) * (size_t)nd);
        int m = orc_bfs_order(nd, order);
        double st[ORC_ARM_DOF];
        for (int i = 0; i < m; ++i) {
            orc_interpolate_d(a, b, (double)order[i] / (double)nd, ORC_ARM_DOF, st);
            ++cnt;
            if (!orc_state_valid_d(env, st, NULL)) {
                result = 0;
                break;
            }
        }
        free(order);
    }
    if (checked) *checked = cnt;
    return result;
}

/* SpaceInformation::checkMotion(states, count) -- base/src/SpaceInformation.cpp:301-341:
 * count 1 -> isValid(front); count >= 2 -> front, back, then a FIFO of (first, last) index pairs
 * seeded with (0, count-1): check mid = (first+last)/2, push (first, mid) if first < mid-1 and
 * (mid, last) if last > mid+1.  *checked = number of isValid calls made. */
int orc_check_states(const orc_env *env, const double *states, int count, int *checked)
{
    int cnt = 0, result = 1;
    if (count == 1) {
        cnt = 1;
        result = orc_state_valid_d(env, states, NULL);
    } else if (count > 1) {
        ++cnt;
        if (!orc_state_valid_d(env, states, NULL)) {
            result = 0;
            goto done;
        }
        ++cnt;
        if (!orc_state_valid_d(env, states + (size_t)(count - 1) * ORC_ARM_DOF, NULL)) {
            result = 0;
            goto done;
        }
        if (count > 2) {
            /* each index pair is pushed at most once per interior index: count slots suffice */
            int *qa = (int *)malloc(sizeof(int) * 2 * (size_t)count);
            int head = 0, tail = 0;
            qa[2 * tail] = 0;
            qa[2 * tail + 1] = count - 1;
            ++tail;
            while (head < tail) {
                const int f = qa[2 * head], l = qa[2 * head + 1];
                const int mid = (f + l) / 2;
                ++cnt;
                if (!orc_state_valid_d(env, states + (size_t)mid * ORC_ARM_DOF, NULL)) {
                    result = 0;
                    break;
                }
                ++head;
                if (f < mid - 1) {
                    qa[2 * tail] = f;
                    qa[2 * tail + 1] = mid;
                    ++tail;
                }
                if (l > mid + 1) {
                    qa[2 * tail] = mid;
                    qa[2 * tail + 1] = l;
                    ++tail;
                }
            }
            free(qa);
        }
    }
done:
    if (checked) *checked = cnt;
    return result;
}

/* SpaceInformation::checkMotion(states, count, firstInvalidStateIndex) -- :288-299: sequential scan. */
int orc_check_states_first_invalid(const orc_env *env, const double *states, int count, int *first_invalid)
{
    for (int i = 0; i < count; ++i)
        if (!orc_state_valid_d(env, states + (size_t)i * ORC_ARM_DOF, NULL)) {
            if (first_invalid) *first_invalid = i;
            return 0;
        }
    return 1;
}

int orc_check_motion_last_valid(const orc_env *env, const double a[ORC_ARM_DOF], const double b[ORC_ARM_DOF],
                                double lvs, unsigned factor, double *last_t, int *checked)
{
    int result = 1, cnt = 0;
    int nd = (int)orc_valid_segment_count(a, b, ORC_ARM_DOF, lvs, factor);
    if (nd > 1) {
        double st[ORC_ARM_DOF];
        for (int j = 1; j < nd; ++j) {
            orc_interpolate_d(a, b, (double)j / (double)nd, ORC_ARM_DOF, st);
            ++cnt;
            if (!orc_state_valid_d(env, st, NULL)) {
                if (last_t) *last_t = (double)(j - 1) / (double)nd;
                result = 0;
                break;
            }
        }
    }
    if (result) {
        ++cnt;
        if (!orc_state_valid_d(env, b, NULL)) {
            if (last_t) *last_t = (double)(nd - 1) / (double)nd;
            result = 0;
        }
    }
    if (checked) *checked = cnt;
    return result;
}

/* ======================================================================
 * std::mt19937 (C++ [rand.eng.mers], parameters of [rand.predef]) and
 * libstdc++'s generate_canonical<double, 53>: two 32-bit draws w1, w2,
 * sum = w1 + w2 * 2^32 (one rounding), u = sum / 2^64, and u is replaced
 * by nextafter(1, 0) when it rounds up to 1.
 * ====================================================================== */
void orc_mt_seed(orc_mt19937 *g, uint32_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < 624; ++i)
        g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + (uint32_t)i;
    g->idx = 624;
}

static void orc_mt_twist(orc_mt19937 *g)
{
    for (int i = 0; i < 624; ++i) {
        uint32_t y = (g->mt[i] & 0x80000000u) | (g->mt[(i + 1) % 624] & 0x7fffffffu);
        uint32_t v = g->mt[(i + 397) % 624] ^ (y >> 1);
        if (y & 1u) v ^= 0x9908b0dfu;
        g->mt[i] = v;
    }
    g->idx = 0;
}

uint32_t orc_mt_next(orc_mt19937 *g)
{
    if (g->idx >= 624) orc_mt_twist(g);
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double orc_canonical(orc_mt19937 *g)
{
    double sum = 0.0;
    double tmp = 1.0;
    for (int k = 0; k < 2; ++k) {
        sum += (double)orc_mt_next(g) * tmp;
        tmp *= 4294967296.0;
    }
    double ret = sum / tmp;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

double orc_uniform_real(orc_mt19937 *g, double lo, double hi)
{
    /* uniform_real_distribution<>(0,1): (1 - 0) * canonical + 0 */
    double u = 1.0 * orc_canonical(g) + 0.0;
    return (hi - lo) * u + lo;
}

void orc_sample_uniform(orc_mt19937 *g, const double *lo, const double *hi, int dim, long n, double *out)
{
    for (long i = 0; i < n; ++i)
        for (int d = 0; d < dim; ++d)
            out[i * dim + d] = orc_uniform_real(g, lo[d], hi[d]);
}

/* UniformValidStateSampler::sample -- base/samplers/src/UniformValidStateSampler.cpp:46-57:
 * do { sampleUniform(state); valid = isValid(state); ++attempts; } while (!valid && attempts < attempts_);
 * n_calls successive calls on one generator; out [n_calls][7] = each call's final state. */
void orc_sample_valid(const orc_env *env, orc_mt19937 *g, const double *lo, const double *hi, long n_calls,
                      int attempts, double *out, uint8_t *found, int *draws)
{
    for (long c = 0; c < n_calls; ++c) {
        double *st = out + c * ORC_ARM_DOF;
        int a = 0, valid = 0;
        do {
            orc_sample_uniform(g, lo, hi, ORC_ARM_DOF, 1, st);
            valid = orc_state_valid_d(env, st, NULL);
            ++a;
        } while (!valid && a < attempts);
        found[c] = (uint8_t)valid;
        if (draws) draws[c] = a;
    }
}

void orc_quaternion(orc_mt19937 *g, double q[4])
{
    const double pi = 3.141592653589793238462643383279502884;
    double x0 = orc_uniform_real(g, 0.0, 1.0);
    double r1 = sqrt(1.0 - x0), r2 = sqrt(x0);
    double t1 = 2.0 * pi * orc_uniform_real(g, 0.0, 1.0);
    double t2 = 2.0 * pi * orc_uniform_real(g, 0.0, 1.0);
    double c1 = cos(t1), s1 = sin(t1);
    double c2 = cos(t2), s2 = sin(t2);
    q[0] = s1 * r1;
    q[1] = c1 * r1;
    q[2] = s2 * r2;
    q[3] = c2 * r2;
}

/* ======================================================================
 * Nearest neighbours, FP32 metric.  Distances are evaluated as the
 * reference's distFun_(element, query): RealVector diff = elem - query,
 * accumulated from 0 in index order, then sqrt; SE3 = translation distance
 * + SO3 arc length (weights 1.0, 1.0), arc = 0 when |dot| > (float)(1-1e-9)
 * (= 1.0f), else acos(|dot|).  Ranking key is (dist, index).
 * ====================================================================== */
float orc_l2_dist(const float *elem, const float *query, int dim)
{
    float dist = 0.0f;
    for (int i = 0; i < dim; ++i) {
        float diff = elem[i] - query[i];
        dist += diff * diff;
    }
    return sqrtf(dist);
}

float orc_se3_dist(const float *e, const float *q)
{
    float dt = orc_l2_dist(e, q, 3);
    float dq = fabsf(((e[3] * q[3] + e[4] * q[4]) + e[5] * q[5]) + e[6] * q[6]);
    float dr = (dq > 1.0f) ? 0.0f : orc_acosf(dq);
    float dist = 0.0f;
    dist += 1.0f * dt;
    dist += 1.0f * dr;
    return dist;
}

static float orc_metric(int metric, const float *e, const float *q, int dim)
{
    return metric == 1 ? orc_se3_dist(e, q) : orc_l2_dist(e, q, dim);
}

typedef struct {
    float d;
    int i;
} orc_pair;

static int pair_less(const orc_pair *a, const orc_pair *b)
{
    return a->d < b->d || (a->d == b->d && a->i < b->i);
}

static int pair_cmp(const void *pa, const void *pb)
{
    const orc_pair *a = (const orc_pair *)pa, *b = (const orc_pair *)pb;
    if (pair_less(a, b)) return -1;
    if (pair_less(b, a)) return 1;
    return 0;
}

void orc_knn(int metric, const float *data, long n, int dim, const float *query, int k, int *out_idx,
             float *out_dist)
{
    /* insertion into a sorted array of size <= k */
    orc_pair *best = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)(k > 0 ? k : 1));
    int m = 0;
    for (long i = 0; i < n; ++i) {
        orc_pair p = {orc_metric(metric, data + i * dim, query, dim), (int)i};
        if (m == k && !pair_less(&p, &best[k - 1])) continue;
        int pos = (m < k) ? m++ : k - 1;
        while (pos > 0 && pair_less(&p, &best[pos - 1])) {
            best[pos] = best[pos - 1];
            --pos;
        }
        best[pos] = p;
    }
    for (int j = 0; j < k; ++j) {
        out_idx[j] = j < m ? best[j].i : -1;
        out_dist[j] = j < m ? best[j].d : INFINITY;
    }
    free(best);
}

long orc_radius(int metric, const float *data, long n, int dim, const float *query, float r, int *out_idx,
                float *out_dist, long cap)
{
    long cnt = 0, alloc = 64;
    orc_pair *v = (orc_pair *)malloc(sizeof(orc_pair) * (size_t)alloc);
    for (long i = 0; i < n; ++i) {
        float d = orc_metric(metric, data + i * dim, query, dim);
        if (d <= r) {
            if (cnt == alloc) {
                alloc *= 2;
                v = (orc_pair *)realloc(v, sizeof(orc_pair) * (size_t)alloc);
            }
            v[cnt].d = d;
            v[cnt].i = (int)i;
            ++cnt;
        }
    }
    qsort(v, (size_t)cnt, sizeof(orc_pair), pair_cmp);
    for (long j = 0; j < cnt && j < cap; ++j) {
        out_idx[j] = v[j].i;
        out_dist[j] = v[j].d;
    }
    free(v);
    return cnt;
}

long orc_nearest(int metric, const float *data, long n, int dim, const float *query)
{
    long pos = -1;
    float dmin = 0.0f;
    for (long i = 0; i < n; ++i) {
        float d = orc_metric(metric, data + i * dim, query, dim);
        if (pos < 0 || dmin > d) {
            pos = i;
            dmin = d;
        }
    }
    return pos;
}

/* ======================================================================
 * Batch drivers over POSIX threads (CPU baseline timing)
 * ====================================================================== */
#include <pthread.h>

typedef struct {
    int kind;
    const orc_env *env;
    const double *a, *b;
    const float *q;
    long n, lo, hi;
    double lvs;
    unsigned factor;
    uint8_t *out;
    int32_t *checked;
    int metric, dim, k;
    const float *data;
    long ndata;
    int *idx;
    float *dist;
} orc_job;

static void *orc_worker(void *p)
{
    orc_job *j = (orc_job *)p;
    for (long i = j->lo; i < j->hi; ++i) {
        if (j->kind == 0) {
            int c = 0;
            j->out[i] = (uint8_t)orc_check_motion(j->env, j->a + i * ORC_ARM_DOF, j->b + i * ORC_ARM_DOF, j->lvs,
                                                  j->factor, &c);
            if (j->checked) j->checked[i] = c;
        } else if (j->kind == 1) {
            float q[ORC_ARM_DOF];
            for (int d = 0; d < ORC_ARM_DOF; ++d) q[d] = j->q[(long)d * j->n + i];
            j->out[i] = (uint8_t)orc_state_valid(j->env, q, NULL);
        } else {
            orc_knn(j->metric, j->data, j->ndata, j->dim, j->q + i * j->dim, j->k, j->idx + i * j->k,
                    j->dist + i * j->k);
        }
    }
    return NULL;
}

static void orc_run(orc_job proto, long n, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    pthread_t th[256];
    orc_job jobs[256];
    long chunk = (n + nthreads - 1) / nthreads;
    int started = 0;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t] = proto;
        jobs[t].lo = t * chunk;
        jobs[t].hi = (t + 1) * chunk < n ? (t + 1) * chunk : n;
        if (jobs[t].lo >= jobs[t].hi) break;
        if (nthreads == 1) {
            orc_worker(&jobs[t]);
        } else {
            pthread_create(&th[t], NULL, orc_worker, &jobs[t]);
            ++started;
        }
    }
    for (int t = 0; t < started; ++t) pthread_join(th[t], NULL);
}

void orc_check_motions_batch(const orc_env *env, const double *a_rows, const double *b_rows, long n, double lvs,
                             unsigned factor, uint8_t *out, int32_t *checked, int nthreads)
{
    orc_job p;
    memset(&p, 0, sizeof(p));
    p.kind = 0;
    p.env = env;
    p.a = a_rows;
    p.b = b_rows;
    p.n = n;
    p.lvs = lvs;
    p.factor = factor;
    p.out = out;
    p.checked = checked;
    orc_run(p, n, nthreads);
}

void orc_validate_states_batch(const orc_env *env, const float *q_soa, long n, uint8_t *out, int nthreads)
{
    orc_job p;
    memset(&p, 0, sizeof(p));
    p.kind = 1;
    p.env = env;
    p.q = q_soa;
    p.n = n;
    p.out = out;
    orc_run(p, n, nthreads);
}

void orc_knn_batch(int metric, const float *data, long n, int dim, const float *queries, long nq, int k,
                   int *out_idx, float *out_dist, int nthreads)
{
    orc_job p;
    memset(&p, 0, sizeof(p));
    p.kind = 2;
    p.metric = metric;
    p.data = data;
    p.ndata = n;
    p.dim = dim;
    p.q = queries;
    p.n = nq;
    p.k = k;
    p.idx = out_idx;
    p.dist = out_dist;
    orc_run(p, nq, nthreads);
}

/* ======================================================================
 * RRT-Connect restatement (see oracle.h)
 * ====================================================================== */
typedef struct {
    double *s;  /* [cap][7] */
    int *parent;
    int n, cap;
} orc_tree;

static void tree_push(orc_tree *t, const double *st, int parent)
{
    if (t->n == t->cap) {
        t->cap = t->cap ? 2 * t->cap : 256;
        t->s = (double *)realloc(t->s, sizeof(double) * ORC_ARM_DOF * (size_t)t->cap);
        t->parent = (int *)realloc(t->parent, sizeof(int) * (size_t)t->cap);
    }
    memcpy(t->s + (size_t)t->n * ORC_ARM_DOF, st, sizeof(double) * ORC_ARM_DOF);
    t->parent[t->n] = parent;
    ++t->n;
}

enum { ORC_TRAPPED = 0, ORC_ADVANCED = 1, ORC_REACHED = 2 };

/* growTree: nearest (first strict minimum of distance(elem, target)), steer
 * by max_distance, check the motion in the direction of tree growth. */
/* growTree (RRTConnect.cpp:116-196).  With addIntermediateStates
 * (:148-183) the motion's getMotionStates (SpaceInformation.cpp:215-286) with
 * count = validSegmentCount(a, b) intermediate states -- count + 1 segments,
 * states a + (b - a) j / (count + 1) -- are added as a chain: from the start
 * tree (a = nearest, b = new state) states 1 .. count + 1, from the goal tree
 * (a = new state, b = nearest) states count .. 0, so the last added is always
 * the new state. */
static int grow_tree(const orc_env *env, orc_tree *t, int start_tree, const double *target, double maxd, double lvs,
                     unsigned factor, int intermediate, int *xmotion)
{
    int nm = 0;
    double dmin = 0.0;
    for (int i = 0; i < t->n; ++i) {
        double d = orc_distance_d(t->s + (size_t)i * ORC_ARM_DOF, target, ORC_ARM_DOF);
        if (i == 0 || dmin > d) {
            nm = i;
            dmin = d;
        }
    }
    double ns[ORC_ARM_DOF];
    memcpy(ns, t->s + (size_t)nm * ORC_ARM_DOF, sizeof(ns)); /* the tree may grow below */
    int reach = 1;
    double x[ORC_ARM_DOF];
    const double *dst = target;
    double d = orc_distance_d(ns, target, ORC_ARM_DOF);
    if (d > maxd) {
        orc_interpolate_d(ns, target, maxd / d, ORC_ARM_DOF, x);
        int equal = 1;
        for (int k = 0; k < ORC_ARM_DOF; ++k)
            if (fabs(ns[k] - x[k]) > 2.220446049250313080847e-16 * 2.0) equal = 0;
        if (equal) return ORC_TRAPPED;
        dst = x;
        reach = 0;
    }
    int valid;
    if (start_tree)
        valid = orc_check_motion(env, ns, dst, lvs, factor, NULL);
    else
        valid = orc_state_valid_d(env, dst, NULL) && orc_check_motion(env, dst, ns, lvs, factor, NULL);
    if (!valid) return ORC_TRAPPED;
    if (intermediate) {
        const double *a = start_tree ? ns : dst, *b = start_tree ? dst : ns;
        const unsigned count = orc_valid_segment_count(a, b, ORC_ARM_DOF, lvs, factor);
        const unsigned seg = count + 1;
        double st[ORC_ARM_DOF];
        int parent = nm;
        for (unsigned i = 1; i <= seg; ++i) {
            /* start tree: state i; goal tree: state seg - i (0 = the new state) */
            const unsigned j = start_tree ? i : seg - i;
            if (j == 0) memcpy(st, a, sizeof(st));
            else if (j == seg) memcpy(st, b, sizeof(st));
            else orc_interpolate_d(a, b, (double)j / (double)seg, ORC_ARM_DOF, st);
            tree_push(t, st, parent);
            parent = t->n - 1;
        }
    } else {
        tree_push(t, dst, nm);
    }
    *xmotion = t->n - 1;
    return reach ? ORC_REACHED : ORC_ADVANCED;
}

/* SpaceInformation::satisfiesBounds on [-pi, pi]^7
 * (RealVectorStateSpace.cpp:200-208: within machine epsilon of the bounds). */
static int in_bounds(const double *q)
{
    const double pi = 3.141592653589793238462643383279502884, eps = 2.220446049250313080847e-16;
    for (int k = 0; k < ORC_ARM_DOF; ++k)
        if (q[k] - eps > pi || q[k] + eps < -pi) return 0;
    return 1;
}

int orc_rrt_connect_ex(const orc_env *env, const double *starts, int n_starts, const double *goals, int n_goals,
                       uint32_t seed, double max_distance, double lvs, unsigned factor, int max_iters,
                       int intermediate, double *path, int path_cap, int *path_len, int *iterations,
                       long *tree_nodes)
{
    const double pi = 3.141592653589793238462643383279502884;
    orc_tree ts = {0}, tg = {0};
    orc_mt19937 rng;
    orc_mt_seed(&rng, seed);
    int start_tree = 1, status = ORC_PLAN_FAILED, it = 0, sampled_goals = 0;
    double rstate[ORC_ARM_DOF];
    *path_len = 0;
    /* PlannerInputStates::nextStart (Planner.cpp:228-256): every start state
     * in order, skipping those out of bounds or invalid */
    for (int i = 0; i < n_starts; ++i) {
        const double *q = starts + (size_t)i * ORC_ARM_DOF;
        if (in_bounds(q) && orc_state_valid_d(env, q, NULL)) tree_push(&ts, q, -1);
    }
    if (ts.n == 0) status = ORC_PLAN_INVALID_START;
    for (it = 0; it < max_iters && ts.n > 0; ++it) {
        orc_tree *tree = start_tree ? &ts : &tg;
        int tgi_start = start_tree;
        start_tree = !start_tree;
        orc_tree *other = start_tree ? &ts : &tg;
        /* RRTConnect.cpp:251-269: while the goal tree is empty, the next valid
         * goal state (nextGoal(ptc) tries them all); later, one goal state per
         * iteration (nextGoal() tries one) while fewer than half the goal
         * tree's size have been sampled (GoalStates: in order, Planner.cpp:
         * 265-336) */
        if (tg.n == 0 || sampled_goals < tg.n / 2) {
            const int empty = tg.n == 0;
            while (sampled_goals < n_goals) {
                const double *q = goals + (size_t)sampled_goals * ORC_ARM_DOF;
                ++sampled_goals;
                if (in_bounds(q) && orc_state_valid_d(env, q, NULL)) {
                    tree_push(&tg, q, -1);
                    break;
                }
                if (!empty) break;
            }
            if (tg.n == 0) {
                status = ORC_PLAN_INVALID_GOAL;
                break;
            }
        }
        for (int k = 0; k < ORC_ARM_DOF; ++k) rstate[k] = orc_uniform_real(&rng, -pi, pi);
        int xmotion = -1;
        int gs = grow_tree(env, tree, tgi_start, rstate, max_distance, lvs, factor, intermediate, &xmotion);
        if (gs == ORC_TRAPPED) continue;
        int added = xmotion;
        if (gs != ORC_REACHED) memcpy(rstate, tree->s + (size_t)added * ORC_ARM_DOF, sizeof(rstate));
        tgi_start = start_tree;
        int gsc = grow_tree(env, other, tgi_start, rstate, max_distance, lvs, factor, intermediate, &xmotion);
        if (gsc == ORC_TRAPPED) tgi_start = !tgi_start;
        while (gsc == ORC_ADVANCED)
            gsc = grow_tree(env, other, tgi_start, rstate, max_distance, lvs, factor, intermediate, &xmotion);
        if (gsc != ORC_REACHED) continue;
        /* tgi.start tells which tree xmotion belongs to; GoalStates accepts
         * every start/goal pair (Goal::isStartGoalPairValid) */
        int sm = tgi_start ? xmotion : added;
        int gm = tgi_start ? added : xmotion;
        if (ts.parent[sm] >= 0) sm = ts.parent[sm];
        else gm = tg.parent[gm];
        int n1 = 0, n2 = 0;
        for (int m = sm; m >= 0; m = ts.parent[m]) ++n1;
        for (int m = gm; m >= 0; m = tg.parent[m]) ++n2;
        *path_len = n1 + n2;
        int pos = n1 - 1;
        for (int m = sm; m >= 0; m = ts.parent[m], --pos)
            if (pos < path_cap) memcpy(path + (size_t)pos * ORC_ARM_DOF, ts.s + (size_t)m * ORC_ARM_DOF, sizeof(rstate));
        pos = n1;
        for (int m = gm; m >= 0; m = tg.parent[m], ++pos)
            if (pos < path_cap) memcpy(path + (size_t)pos * ORC_ARM_DOF, tg.s + (size_t)m * ORC_ARM_DOF, sizeof(rstate));
        status = ORC_PLAN_SOLVED;
        ++it;
        break;
    }
    *iterations = it;
    if (tree_nodes) *tree_nodes = (long)ts.n + tg.n;
    free(ts.s);
    free(ts.parent);
    free(tg.s);
    free(tg.parent);
    return status;
}

int orc_rrt_connect(const orc_env *env, const double start[ORC_ARM_DOF], const double goal[ORC_ARM_DOF],
                    uint32_t seed, double max_distance, double lvs, unsigned factor, int max_iters, double *path,
                    int path_cap, int *path_len, int *iterations, long *tree_nodes)
{
    return orc_rrt_connect_ex(env, start, 1, goal, 1, seed, max_distance, lvs, factor, max_iters, 0, path, path_cap,
                              path_len, iterations, tree_nodes) == ORC_PLAN_SOLVED;
}
