// CPU restatement of ompl::NearestNeighborsGNAT (datastructures/
// NearestNeighborsGNAT.h:72-803) with GreedyKCenters pivot selection
// (datastructures/GreedyKCenters.h:82-117), without Eigen or Boost.  Test
// infrastructure: it is the nearest-neighbour CPU baseline the GPU search is
// timed against (GNAT is what SelfConfig::getDefaultNearestNeighbors gives
// planners, tools/config/SelfConfig.h:105-116), and its results are checked
// against the Linear restatement the way the reference's own test checks
// GNAT (tests/datastructures/nearestneighbors.cpp: equal distance sets).
//
// Followed from the reference:
//   * bulk add: the first element is the root's pivot, the rest its data,
//     then split while a leaf holds more than max(maxNumPtsPerLeaf, degree)
//     (:161-176, :494-498);
//   * split: `degree` pivots by greedy k-centers (first centre drawn
//     uniformly, then repeatedly the point farthest from the centres so
//     far, stopping early when that distance is below DBL_EPSILON); each
//     point goes to its nearest pivot (first minimum), updating the child's
//     radius and every child's range for the receiving child; children get
//     degree clamp(degree * |child| / |parent|, minDegree, maxDegree)
//     (:500-541, GreedyKCenters.h:82-117);
//   * kNN: a bounded max-queue of (distance, element) and a node queue
//     ordered by distance-to-pivot minus maxRadius; a node is visited unless
//     the queue is full and its pivot distance is outside
//     [minRadius - d_k, maxRadius + d_k]; children are scanned in the
//     rotating order (i + offset) % size, pruning siblings by the ranges
//     (:335-356, :565-612); radius search alike with d_k = r (:358-376,
//     :614-662); results ascending (:379-384).
// Defaults: degree 8, minDegree 4, maxDegree 12, maxNumPtsPerLeaf 50
// (:95-114).  Distances are FP64 over the FP32 coordinates (metric 0: L2;
// metric 1: SE3 = |t1 - t2| + acos(|q1 . q2|), 0 when |q1 . q2| > 1 - 1e-9,
// base/spaces/src/SO3StateSpace.cpp:253-262).  Equal distances are ordered
// by element index.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <queue>
#include <random>
#include <thread>
#include <utility>
#include <vector>

#include "oracle.h"

namespace {

struct Gnat;

struct Node {
    int degree;
    int32_t pivot;
    double min_radius = std::numeric_limits<double>::infinity();
    double max_radius = -std::numeric_limits<double>::infinity();
    std::vector<double> min_range, max_range;
    std::vector<int32_t> data;
    std::vector<Node *> children;

    Node(int deg, int32_t piv)
        : degree(deg), pivot(piv), min_range(deg, std::numeric_limits<double>::infinity()),
          max_range(deg, -std::numeric_limits<double>::infinity())
    {
    }
    ~Node()
    {
        for (Node *c : children) delete c;
    }
    void update_radius(double d)
    {
        if (min_radius > d) min_radius = d;
        if (max_radius < d) max_radius = d;
    }
    void update_range(int i, double d)
    {
        if (min_range[i] > d) min_range[i] = d;
        if (max_range[i] < d) max_range[i] = d;
    }
};

struct Gnat {
    int metric, dim;
    const float *pts;
    int64_t n;
    int degree = 8, min_degree = 4, max_degree = 12, leaf = 50;
    std::mt19937 rng;
    Node *root = nullptr;

    double dist(int32_t a, const float *q) const
    {
        const float *p = pts + (int64_t)a * dim;
        if (metric == 0) {
            double s = 0.0;
            for (int k = 0; k < dim; ++k) {
                const double d = (double)p[k] - (double)q[k];
                s += d * d;
            }
            return std::sqrt(s);
        }
        double s = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = (double)p[k] - (double)q[k];
            s += d * d;
        }
        double dot = 0.0;
        for (int k = 3; k < 7; ++k) dot += (double)p[k] * (double)q[k];
        dot = std::fabs(dot);
        const double rot = dot > 1.0 - 1e-9 ? 0.0 : std::acos(dot);
        return std::sqrt(s) + rot;
    }
    double dist(int32_t a, int32_t b) const { return dist(a, pts + (int64_t)b * dim); }

    bool need_split(const Node *nd) const
    {
        const size_t sz = nd->data.size();
        return sz > (size_t)leaf && sz > (size_t)nd->degree;
    }

    // GreedyKCenters::kcenters over nd->data: centres (indices into data) and
    // dists[j * k + i] = distance(data[j], data[centre i])
    void kcenters(const std::vector<int32_t> &data, int k, std::vector<int> &centres, std::vector<double> &dists)
    {
        const size_t m = data.size();
        std::vector<double> min_dist(m, std::numeric_limits<double>::infinity());
        dists.assign(m * (size_t)k, 0.0);
        centres.clear();
        centres.push_back(std::uniform_int_distribution<int>(0, (int)m - 1)(rng));
        for (int i = 1; i < k; ++i) {
            int ind = 0;
            const int32_t c = data[(size_t)centres[(size_t)i - 1]];
            double max_d = -std::numeric_limits<double>::infinity();
            for (size_t j = 0; j < m; ++j) {
                const double d = dist(data[j], c);
                dists[j * (size_t)k + (size_t)i - 1] = d;
                if (d < min_dist[j]) min_dist[j] = d;
                if (min_dist[j] > max_d) {
                    ind = (int)j;
                    max_d = min_dist[j];
                }
            }
            if (max_d < std::numeric_limits<double>::epsilon()) break;
            centres.push_back(ind);
        }
        const int32_t c = data[(size_t)centres.back()];
        const size_t i = centres.size() - 1;
        for (size_t j = 0; j < m; ++j) dists[j * (size_t)k + i] = dist(data[j], c);
    }

    void split(Node *nd)
    {
        std::vector<int> piv;
        std::vector<double> dd;
        const int k0 = nd->degree;
        kcenters(nd->data, k0, piv, dd);
        // a child's ranges hold one entry per sibling: sized to this node's degree
        for (int p : piv) nd->children.push_back(new Node(k0, nd->data[(size_t)p]));
        const int deg = (int)piv.size();
        nd->degree = deg;
        for (size_t j = 0; j < nd->data.size(); ++j) {
            int k = 0;
            for (int i = 1; i < deg; ++i)
                if (dd[j * (size_t)k0 + (size_t)i] < dd[j * (size_t)k0 + (size_t)k]) k = i;
            Node *child = nd->children[(size_t)k];
            if ((int)j != piv[(size_t)k]) {
                child->data.push_back(nd->data[j]);
                child->update_radius(dd[j * (size_t)k0 + (size_t)k]);
            }
            for (int i = 0; i < deg; ++i) nd->children[(size_t)i]->update_range(k, dd[j * (size_t)k0 + (size_t)i]);
        }
        const size_t total = nd->data.size();
        for (Node *c : nd->children) {
            c->degree = std::min(std::max((int)(((size_t)deg * c->data.size()) / total), min_degree), max_degree);
            if (c->min_radius >= std::numeric_limits<double>::infinity()) c->min_radius = c->max_radius = 0.0;
        }
        std::vector<int32_t>().swap(nd->data);
        for (Node *c : nd->children)
            if (need_split(c)) split(c);
    }
    using NearQueue = std::priority_queue<std::pair<double, int32_t>>;
    struct NodeDist {
        const Node *node;
        double d;
    };
    struct NodeDistCompare {
        bool operator()(const NodeDist &a, const NodeDist &b) const
        {
            return (a.d - a.node->max_radius) > (b.d - b.node->max_radius);
        }
    };
    using NodeQueue = std::priority_queue<NodeDist, std::vector<NodeDist>, NodeDistCompare>;

    static bool insert_k(NearQueue &nbh, size_t k, int32_t e, double d)
    {
        if (nbh.size() < k) {
            nbh.emplace(d, e);
            return true;
        }
        if (std::make_pair(d, e) < nbh.top()) {
            nbh.pop();
            nbh.emplace(d, e);
            return true;
        }
        return false;
    }

    void node_knn(const Node *nd, const float *q, size_t k, NearQueue &nbh, NodeQueue &nq, size_t &offset) const
    {
        for (int32_t e : nd->data) insert_k(nbh, k, e, dist(e, q));
        if (nd->children.empty()) return;
        const size_t sz = nd->children.size(), off = offset++;
        std::vector<double> dp(sz);
        std::vector<int> perm(sz);
        for (size_t i = 0; i < sz; ++i) perm[i] = (int)((i + off) % sz);
        for (size_t i = 0; i < sz; ++i)
            if (perm[i] >= 0) {
                const Node *c = nd->children[(size_t)perm[i]];
                dp[(size_t)perm[i]] = dist(c->pivot, q);
                insert_k(nbh, k, c->pivot, dp[(size_t)perm[i]]);
                if (nbh.size() == k) {
                    const double dk = nbh.top().first;
                    for (size_t j = 0; j < sz; ++j)
                        if (perm[j] >= 0 && i != j &&
                            (dp[(size_t)perm[i]] - dk > c->max_range[(size_t)perm[j]] ||
                             dp[(size_t)perm[i]] + dk < c->min_range[(size_t)perm[j]]))
                            perm[j] = -1;
                }
            }
        const double dk = nbh.top().first;
        for (int p : perm)
            if (p >= 0) {
                const Node *c = nd->children[(size_t)p];
                if (nbh.size() < k || (dp[(size_t)p] - dk <= c->max_radius && dp[(size_t)p] + dk >= c->min_radius))
                    nq.push(NodeDist{c, dp[(size_t)p]});
            }
    }

    void knn(const float *q, size_t k, int32_t *idx, double *dd, size_t &offset) const
    {
        NearQueue nbh;
        NodeQueue nq;
        insert_k(nbh, k, root->pivot, dist(root->pivot, q));
        node_knn(root, q, k, nbh, nq, offset);
        while (!nq.empty()) {
            const double dk = nbh.top().first;
            const NodeDist nd = nq.top();
            nq.pop();
            if (nbh.size() == k && (nd.d > nd.node->max_radius + dk || nd.d < nd.node->min_radius - dk)) continue;
            node_knn(nd.node, q, k, nbh, nq, offset);
        }
        const size_t got = nbh.size();
        for (size_t i = 0; i < k; ++i) {
            idx[i] = -1;
            dd[i] = std::numeric_limits<double>::infinity();
        }
        for (size_t i = got; i-- > 0; nbh.pop()) {
            idx[i] = nbh.top().second;
            dd[i] = nbh.top().first;
        }
    }

    void node_radius(const Node *nd, const float *q, double r, std::vector<std::pair<double, int32_t>> &out,
                     NodeQueue &nq, size_t &offset) const
    {
        for (int32_t e : nd->data) {
            const double d = dist(e, q);
            if (d <= r) out.emplace_back(d, e);
        }
        if (nd->children.empty()) return;
        const size_t sz = nd->children.size(), off = offset++;
        std::vector<double> dp(sz);
        std::vector<int> perm(sz);
        for (size_t i = 0; i < sz; ++i) perm[i] = (int)((i + off) % sz);
        for (size_t i = 0; i < sz; ++i)
            if (perm[i] >= 0) {
                const Node *c = nd->children[(size_t)perm[i]];
                dp[(size_t)perm[i]] = dist(c->pivot, q);
                if (dp[(size_t)perm[i]] <= r) out.emplace_back(dp[(size_t)perm[i]], c->pivot);
                for (size_t j = 0; j < sz; ++j)
                    if (perm[j] >= 0 && i != j &&
                        (dp[(size_t)perm[i]] - r > c->max_range[(size_t)perm[j]] ||
                         dp[(size_t)perm[i]] + r < c->min_range[(size_t)perm[j]]))
                        perm[j] = -1;
            }
        for (int p : perm)
            if (p >= 0) {
                const Node *c = nd->children[(size_t)p];
                if (dp[(size_t)p] - r <= c->max_radius && dp[(size_t)p] + r >= c->min_radius)
                    nq.push(NodeDist{c, dp[(size_t)p]});
            }
    }

    void radius(const float *q, double r, std::vector<std::pair<double, int32_t>> &out, size_t &offset) const
    {
        out.clear();
        NodeQueue nq;
        const double d0 = dist(root->pivot, q);
        if (d0 <= r) out.emplace_back(d0, root->pivot);
        node_radius(root, q, r, out, nq, offset);
        while (!nq.empty()) {
            const NodeDist nd = nq.top();
            nq.pop();
            if (nd.d > nd.node->max_radius + r || nd.d < nd.node->min_radius - r) continue;
            node_radius(nd.node, q, r, out, nq, offset);
        }
        std::sort(out.begin(), out.end());
    }
};

template <class F>
void parallel_for(int64_t n, int threads, F f)
{
    if (threads <= 1 || n < 2) {
        f(0, n, 0);
        return;
    }
    std::vector<std::thread> th;
    const int64_t per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t b = t * per, e = std::min(n, b + per);
        if (b >= e) break;
        th.emplace_back([=] { f(b, e, t); });
    }
    for (auto &x : th) x.join();
}

} // namespace

struct orc_gnat {
    Gnat g;
};

extern "C" {

orc_gnat *orc_gnat_build(int metric, const float *pts, int64_t n, int dim, uint32_t seed)
{
    if (n <= 0 || (metric == 1 && dim != 7)) return nullptr;
    auto *h = new orc_gnat();
    Gnat &g = h->g;
    g.metric = metric;
    g.dim = dim;
    g.pts = pts;
    g.n = n;
    g.rng.seed(seed);
    g.root = new Node(g.degree, 0);
    g.root->data.reserve((size_t)n - 1);
    for (int64_t i = 1; i < n; ++i) g.root->data.push_back((int32_t)i);
    if (g.need_split(g.root)) g.split(g.root);
    return h;
}

void orc_gnat_free(orc_gnat *h)
{
    if (!h) return;
    delete h->g.root;
    delete h;
}

void orc_gnat_knn(const orc_gnat *h, const float *q, int64_t nq, int k, int32_t *idx, double *dist, int threads)
{
    const Gnat &g = h->g;
    parallel_for(nq, threads, [&](int64_t b, int64_t e, int t) {
        size_t offset = (size_t)t;
        for (int64_t i = b; i < e; ++i) g.knn(q + i * g.dim, (size_t)k, idx + i * k, dist + i * k, offset);
    });
}

int64_t orc_gnat_radius(const orc_gnat *h, const float *q, int64_t nq, double r, int64_t *counts, int32_t *idx,
                        int64_t cap, int threads)
{
    const Gnat &g = h->g;
    std::vector<std::vector<std::pair<double, int32_t>>> res((size_t)nq);
    parallel_for(nq, threads, [&](int64_t b, int64_t e, int t) {
        size_t offset = (size_t)t;
        for (int64_t i = b; i < e; ++i) g.radius(q + i * g.dim, r, res[(size_t)i], offset);
    });
    int64_t total = 0;
    for (int64_t i = 0; i < nq; ++i) {
        if (counts) counts[i] = (int64_t)res[(size_t)i].size();
        for (const auto &p : res[(size_t)i]) {
            if (idx && total < cap) idx[total] = p.second;
            ++total;
        }
    }
    return total;
}

} // extern "C"
