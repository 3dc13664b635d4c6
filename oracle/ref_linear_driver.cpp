// Runs the reference's own ompl::NearestNeighborsLinear (compiled unchanged
// from datastructures/NearestNeighborsLinear.h of the reference tree) on
// element indices with an FP64 distance function, to pin the FP32 oracle's
// k-NN / radius results.  Checker only; never part of the product.
//
// Distances restate the reference metrics in FP64 over FP32 inputs widened to
// double (what an OMPL state holding the same values computes):
//   metric 0: RealVectorStateSpace::distance  base/spaces/src/RealVectorStateSpace.cpp:230-242
//   metric 1: SE3 = R3 distance + SO3 arcLength, weights 1.0
//             base/src/StateSpace.cpp:1074-1082, base/spaces/src/SO3StateSpace.cpp:253-262
//
// usage: ref_linear <metric> <dim> <data.f32> <n> <queries.f32> <nq> <k> <radius> <out.txt>
// out.txt: per query "K i0 d0 i1 d1 ..." then "R cnt i0 d0 ..." (distances %.17g)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "ompl/datastructures/NearestNeighborsLinear.h"

static std::vector<float> load(const char *path, size_t count)
{
    std::vector<float> v(count);
    FILE *f = std::fopen(path, "rb");
    if (!f || std::fread(v.data(), sizeof(float), count, f) != count) {
        std::fprintf(stderr, "cannot read %s\n", path);
        std::exit(2);
    }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc != 10) {
        std::fprintf(stderr, "usage: ref_linear metric dim data n queries nq k radius out\n");
        return 2;
    }
    int metric = std::atoi(argv[1]);
    int dim = std::atoi(argv[2]);
    long n = std::atol(argv[4]);
    long nq = std::atol(argv[6]);
    int k = std::atoi(argv[7]);
    double radius = std::atof(argv[8]);
    std::vector<float> data = load(argv[3], (size_t)n * dim);
    std::vector<float> qs = load(argv[5], (size_t)nq * dim);
    // element ids >= 0 index the data; id -1 - j denotes query j
    auto coords = [&](long id) -> const float * {
        return id >= 0 ? &data[(size_t)id * dim] : &qs[(size_t)(-1 - id) * dim];
    };
    auto dist = [&](const long &a, const long &b) -> double {
        const float *x = coords(a), *y = coords(b);
        if (metric == 0) {
            double s = 0.0;
            for (int i = 0; i < dim; ++i) {
                double diff = (double)x[i] - (double)y[i];
                s += diff * diff;
            }
            return std::sqrt(s);
        }
        double s = 0.0;
        for (int i = 0; i < 3; ++i) {
            double diff = (double)x[i] - (double)y[i];
            s += diff * diff;
        }
        double dt = std::sqrt(s);
        double dq = std::fabs((double)x[3] * y[3] + (double)x[4] * y[4] + (double)x[5] * y[5] + (double)x[6] * y[6]);
        double dr = dq > 1.0 - 1e-9 ? 0.0 : std::acos(dq);
        double d = 0.0;
        d += 1.0 * dt;
        d += 1.0 * dr;
        return d;
    };
    ompl::NearestNeighborsLinear<long> nn;
    nn.setDistanceFunction(dist);
    std::vector<long> ids((size_t)n);
    for (long i = 0; i < n; ++i) ids[(size_t)i] = i;
    nn.add(ids);
    FILE *out = std::fopen(argv[9], "w");
    std::vector<long> nbh;
    for (long j = 0; j < nq; ++j) {
        long qid = -1 - j;
        nn.nearestK(qid, (size_t)k, nbh);
        std::fprintf(out, "K");
        for (long e : nbh) std::fprintf(out, " %ld %.17g", e, dist(e, qid));
        std::fprintf(out, "\n");
        nn.nearestR(qid, radius, nbh);
        std::fprintf(out, "R %zu", nbh.size());
        for (long e : nbh) std::fprintf(out, " %ld %.17g", e, dist(e, qid));
        std::fprintf(out, "\n");
        long nearest = nn.nearest(qid);
        std::fprintf(out, "N %ld\n", nearest);
    }
    std::fclose(out);
    return 0;
}
