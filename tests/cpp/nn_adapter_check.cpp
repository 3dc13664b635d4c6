// Compile-and-run check of ompl_amd::NearestNeighborsGPU<_T> against the
// reference's own ompl::NearestNeighbors<_T> interface header.  Without a GPU
// it only exercises construction failure paths; with one it compares the
// adapter with NearestNeighborsLinear on a small set.
#include <algorithm>
#include <cmath>
#include <iterator>
#include <cstdio>
#include <memory>
#include <vector>

#include <ompl/datastructures/NearestNeighborsLinear.h>

#include "ompl_amd/nearest_neighbors_gpu.hpp"

struct Pt {
    float c[3];
    int id;
    bool operator==(const Pt &o) const { return id == o.id; }
};

int main(int argc, char **argv)
{
    const bool run_gpu = argc > 1;
    std::unique_ptr<ompl::NearestNeighbors<Pt>> base;
    try {
        base = std::make_unique<ompl_amd::NearestNeighborsGPU<Pt>>(OMPL_AMD_METRIC_L2, 3, [](const Pt &p, float *o) {
            o[0] = p.c[0];
            o[1] = p.c[1];
            o[2] = p.c[2];
        });
    } catch (const ompl::Exception &e) {
        std::printf("construction failed as expected without a device: %s\n", e.what());
        return run_gpu ? 1 : 0;
    }
    if (!run_gpu) return 0;
    auto dist = [](const Pt &a, const Pt &b) {
        double s = 0;
        for (int i = 0; i < 3; ++i) s += ((double)a.c[i] - b.c[i]) * ((double)a.c[i] - b.c[i]);
        return std::sqrt(s);
    };
    base->setDistanceFunction(dist);
    ompl::NearestNeighborsLinear<Pt> lin;
    lin.setDistanceFunction(dist);
    std::vector<Pt> pts;
    unsigned s = 12345;
    for (int i = 0; i < 500; ++i) {
        Pt p;
        for (float &v : p.c) {
            s = s * 1664525u + 1013904223u;
            v = (float)((s >> 8) % 10000) / 1000.0f;
        }
        p.id = i;
        pts.push_back(p);
    }
    base->add(pts);
    lin.add(pts);
    // Index sets must be equal; the only differences allowed are points whose
    // FP64 distance lies within 1e-5 of the boundary (the k-th distance, or the
    // radius), where the GPU's FP32 ranking and Linear's FP64 partial_sort
    // (which leaves ties in no defined order) may legitimately disagree.
    int bad = 0, boundary = 0;
    auto ids = [](const std::vector<Pt> &v) {
        std::vector<int> r;
        for (const Pt &p : v) r.push_back(p.id);
        std::sort(r.begin(), r.end());
        return r;
    };
    auto sym_diff = [&](const std::vector<Pt> &a, const std::vector<Pt> &b) {
        std::vector<int> ia = ids(a), ib = ids(b), d;
        std::set_symmetric_difference(ia.begin(), ia.end(), ib.begin(), ib.end(), std::back_inserter(d));
        return d;
    };
    for (int q = 0; q < 50; ++q) {
        const Pt &qp = pts[(size_t)q * 7];
        std::vector<Pt> a, b;
        base->nearestK(qp, 8, a);
        lin.nearestK(qp, 8, b);
        if (a.size() != b.size()) ++bad;
        const double dk = dist(b.back(), qp);
        for (int id : sym_diff(a, b)) {
            if (std::fabs(dist(pts[(size_t)id], qp) - dk) > 1e-5) ++bad;
            else ++boundary;
        }
        for (size_t j = 0; j < a.size() && j < b.size(); ++j) // ascending order, ranks equal up to ties
            if (a[j].id != b[j].id && std::fabs(dist(a[j], qp) - dist(b[j], qp)) > 1e-5) ++bad;
        base->nearestR(qp, 1.0, a);
        lin.nearestR(qp, 1.0, b);
        for (int id : sym_diff(a, b)) {
            if (std::fabs(dist(pts[(size_t)id], qp) - 1.0) > 1e-5) ++bad;
            else ++boundary;
        }
        if (base->nearest(qp).id != lin.nearest(qp).id) ++bad;
    }
    // k above the device limit: an error, not a silently shorter answer
    bool threw = false;
    try {
        std::vector<Pt> a;
        base->nearestK(pts[0], ompl_amd::NearestNeighborsGPU<Pt>::kMaxK + 1, a);
    } catch (const ompl::Exception &) {
        threw = true;
    }
    if (!threw) ++bad;
    // k above size(): min(k, size()) neighbours, as NearestNeighborsLinear
    {
        ompl_amd::NearestNeighborsGPU<Pt> small(OMPL_AMD_METRIC_L2, 3, [](const Pt &p, float *o) {
            o[0] = p.c[0];
            o[1] = p.c[1];
            o[2] = p.c[2];
        });
        small.setDistanceFunction(dist);
        small.add(std::vector<Pt>(pts.begin(), pts.begin() + 5));
        std::vector<Pt> a;
        small.nearestK(pts[0], 100, a);
        if (a.size() != 5 || a[0].id != 0) ++bad;
    }
    // packed forms: the same lists as one query at a time, id for id
    {
        auto *gpu = static_cast<ompl_amd::NearestNeighborsGPU<Pt> *>(base.get());
        std::vector<Pt> qs;
        std::vector<std::size_t> ks;
        std::vector<double> rs;
        for (int q = 0; q < 40; ++q) {
            qs.push_back(pts[(size_t)q * 11 + 3]);
            ks.push_back((std::size_t)(q % 13));  // 0 .. 12
            rs.push_back(0.4 + 0.05 * (q % 9));
        }
        std::vector<std::vector<Pt>> pk, pr;
        gpu->nearestKPacked(qs, ks, pk);
        gpu->nearestRPacked(qs, rs, pr);
        for (size_t i = 0; i < qs.size(); ++i) {
            std::vector<Pt> a, b;
            base->nearestK(qs[i], ks[i], a);
            base->nearestR(qs[i], rs[i], b);
            if (ids(a) != ids(pk[i]) || a.size() != pk[i].size()) ++bad;
            for (size_t j = 0; j < a.size() && j < pk[i].size(); ++j)
                if (a[j].id != pk[i][j].id) ++bad;
            if (b.size() != pr[i].size()) ++bad;
            for (size_t j = 0; j < b.size() && j < pr[i].size(); ++j)
                if (b[j].id != pr[i][j].id) ++bad;
        }
    }
    base->remove(pts[3]);
    if (base->size() != 499) ++bad;
    std::printf("adapter check: %d mismatches, %d boundary ties\n", bad, boundary);
    return bad ? 1 : 0;
}
