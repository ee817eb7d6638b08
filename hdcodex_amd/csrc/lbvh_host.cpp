// Host parts of the device BVH builder; see lbvh_host.hpp.
#include "lbvh_host.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace hdcx {
namespace {

using lbvh::Box4;
using lbvh::FrontierItem;

double Area(const float lo[3], const float hi[3])
{
    const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
    return dx * dy + dy * dz + dz * dx;
}

void Grow(float lo[3], float hi[3], const float blo[3], const float bhi[3])
{
    for (int x = 0; x < 3; ++x) {
        lo[x] = std::min(lo[x], blo[x]);
        hi[x] = std::max(hi[x], bhi[x]);
    }
}

float Centre(const FrontierItem& f, int axis) { return 0.5F * (f.box.lo[axis] + f.box.hi[axis]); }

constexpr std::uint32_t kMaxSahDepth = 32;  // SAH splits above this depth, median splits below

struct SerialAtomics {
    static std::uint32_t Arrive(std::uint32_t* c) { return (*c)++; }
    static std::uint32_t Append(std::uint32_t* c) { return (*c)++; }
    static void Max(std::uint32_t* c, std::uint32_t v) { *c = std::max(*c, v); }
    static void Flag(std::uint32_t* c) { *c = 1U; }
};

} // namespace

std::uint32_t BuildSahTop(const std::vector<FrontierItem>& items, std::uint32_t firstId,
                          std::vector<std::uint32_t>& child2, std::vector<std::uint32_t>& range2,
                          std::vector<Box4>& box)
{
    const auto count = static_cast<std::uint32_t>(items.size());
    if (count < 2) throw std::invalid_argument("SAH top needs two frontier items");
    child2.clear();
    range2.clear();
    box.clear();
    // Item indices sorted by box centre along each axis (ties by index).
    // Every node owns the same range [b, e) of all three lists; a split
    // keeps each list sorted by partitioning it stably.
    std::vector<std::uint32_t> order[3];
    for (int a = 0; a < 3; ++a) {
        std::vector<std::pair<float, std::uint32_t>> keys(count);
        for (std::uint32_t i = 0; i < count; ++i) keys[i] = {Centre(items[i], a), i};
        std::sort(keys.begin(), keys.end());
        order[a].resize(count);
        for (std::uint32_t i = 0; i < count; ++i) order[a][i] = keys[i].second;
    }
    std::vector<double> rightArea(count), rightCnt(count);
    std::vector<std::uint8_t> left(count);
    std::vector<std::uint32_t> scratch(count);
    // Work items: [b, e) of the lists, the (node, side) slot its reference
    // goes into (node == kNone for the root) and its depth. Left children
    // are taken first, so node ids are in depth-first pre-order.
    struct Work { std::uint32_t b, e, node, side, depth; };
    std::vector<Work> work{{0, count, lbvh::kNone, 0, 0}};
    std::uint32_t root = lbvh::kNone;
    while (!work.empty()) {
        const Work w = work.back();
        work.pop_back();
        std::uint32_t ref;
        if (w.e - w.b == 1) {
            ref = items[order[0][w.b]].ref;
        } else {
            float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for (std::uint32_t j = w.b; j < w.e; ++j) Grow(lo, hi, items[order[0][j]].box.lo, items[order[0][j]].box.hi);
            // Full sweep: every split position along every axis; cost =
            // area(left) * triangles(left) + area(right) * triangles(right);
            // equal costs go to the more even split (coincident items).
            // Below kMaxSahDepth the split is the median along the axis of
            // largest centre extent, which bounds the depth (and the
            // traversal stack) for chains of nested items.
            double best = std::numeric_limits<double>::infinity();
            int bestAxis = 0;
            std::uint32_t mid = w.b + (w.e - w.b) / 2;
            const std::uint32_t half = mid;
            if (w.depth >= kMaxSahDepth) {
                float extent = -1.0F;
                for (int a = 0; a < 3; ++a) {
                    const float ext = Centre(items[order[a][w.e - 1]], a) - Centre(items[order[a][w.b]], a);
                    if (ext > extent) {
                        extent = ext;
                        bestAxis = a;
                    }
                }
            }
            for (int a = 0; a < 3 && w.depth < kMaxSahDepth; ++a) {
                const std::uint32_t* o = order[a].data();
                float alo[3] = {INFINITY, INFINITY, INFINITY}, ahi[3] = {-INFINITY, -INFINITY, -INFINITY};
                double acc = 0.0;
                for (std::uint32_t j = w.e; j-- > w.b + 1;) {
                    const FrontierItem& f = items[o[j]];
                    Grow(alo, ahi, f.box.lo, f.box.hi);
                    acc += f.size;
                    rightArea[j] = Area(alo, ahi);
                    rightCnt[j] = acc;
                }
                for (int x = 0; x < 3; ++x) {
                    alo[x] = INFINITY;
                    ahi[x] = -INFINITY;
                }
                acc = 0.0;
                for (std::uint32_t j = w.b + 1; j < w.e; ++j) {
                    const FrontierItem& f = items[o[j - 1]];
                    Grow(alo, ahi, f.box.lo, f.box.hi);
                    acc += f.size;
                    const double cost = Area(alo, ahi) * acc + rightArea[j] * rightCnt[j];
                    const auto skew = [half](std::uint32_t x) { return x > half ? x - half : half - x; };
                    if (cost < best || (cost == best && skew(j) < skew(mid))) {
                        best = cost;
                        bestAxis = a;
                        mid = j;
                    }
                }
            }
            for (std::uint32_t j = w.b; j < w.e; ++j) left[order[bestAxis][j]] = j < mid ? 1U : 0U;
            for (int a = 0; a < 3; ++a) {
                if (a == bestAxis) continue;
                std::uint32_t* o = order[a].data();
                std::uint32_t l = w.b, r = 0;
                for (std::uint32_t j = w.b; j < w.e; ++j) {
                    if (left[o[j]]) o[l++] = o[j];
                    else scratch[r++] = o[j];
                }
                std::copy(scratch.begin(), scratch.begin() + r, o + l);
            }
            const auto self = static_cast<std::uint32_t>(box.size());
            child2.push_back(lbvh::kNone);
            child2.push_back(lbvh::kNone);
            range2.push_back(0U);
            range2.push_back(lbvh::kTopRangeHi);
            Box4 b{};
            for (int x = 0; x < 3; ++x) {
                b.lo[x] = lo[x];
                b.hi[x] = hi[x];
            }
            box.push_back(b);
            ref = firstId + self;
            work.push_back({mid, w.e, self, 1, w.depth + 1});
            work.push_back({w.b, mid, self, 0, w.depth + 1});
        }
        if (w.node == lbvh::kNone) root = ref;
        else child2[2 * w.node + w.side] = ref;
    }
    return root;
}

Bvh4 BuildBvh4OverBoxes(const std::vector<Box4>& boxes)
{
    if (boxes.empty()) throw std::invalid_argument("no boxes");
    Bvh4 out;
    const auto n = static_cast<std::uint32_t>(boxes.size());
    if (n == 1) {
        Bvh4Node node{};
        for (int i = 0; i < 4; ++i) {
            node.lox[i] = node.loy[i] = node.loz[i] = INFINITY;
            node.hix[i] = node.hiy[i] = node.hiz[i] = -INFINITY;
            node.child[i] = kEmptyChild;
        }
        node.lox[0] = lbvh::PadDown(boxes[0].lo[0]);
        node.loy[0] = lbvh::PadDown(boxes[0].lo[1]);
        node.loz[0] = lbvh::PadDown(boxes[0].lo[2]);
        node.hix[0] = lbvh::PadUp(boxes[0].hi[0]);
        node.hiy[0] = lbvh::PadUp(boxes[0].hi[1]);
        node.hiz[0] = lbvh::PadUp(boxes[0].hi[2]);
        node.child[0] = EncodeLeaf(0, 1);
        out.nodes.push_back(node);
        out.depth = 1;
        return out;
    }
    std::vector<FrontierItem> items(n);
    for (std::uint32_t k = 0; k < n; ++k) items[k] = FrontierItem{k | lbvh::kLeafRef, 1U, 0U, 0U, boxes[k]};
    std::vector<std::uint32_t> child2, range2;
    std::vector<Box4> nodeBox;
    const std::uint32_t root = BuildSahTop(items, 0, child2, range2, nodeBox);
    std::vector<std::uint32_t> ids(n);
    for (std::uint32_t k = 0; k < n; ++k) ids[k] = k;
    const std::uint32_t inner = static_cast<std::uint32_t>(nodeBox.size());
    out.nodes.resize(inner);
    std::vector<std::uint32_t> source(inner, 0U), nodeStack(inner, 0U);
    std::uint32_t counters[4] = {1U, 0U, 0U, 0U};
    source[0] = root;
    std::uint32_t begin = 0, end = 1, level = 0;
    while (begin < end) {
        for (std::uint32_t id = begin; id < end; ++id)
            lbvh::CollapseNode<SerialAtomics>(id, level, source.data(), child2.data(), range2.data(), nodeBox.data(),
                                              boxes.data(), ids.data(), nullptr, nodeStack.data(), counters, inner,
                                              out.nodes.data());
        if (counters[3]) throw std::runtime_error("instance tree collapse overflowed its node pool");
        begin = end;
        end = counters[0];
        ++level;
    }
    out.nodes.resize(counters[0]);
    out.maxStack = counters[1];
    out.depth = counters[2];
    return out;
}

void VertexBounds(const float* positions, std::uint32_t vertexCount, float lo[3], float hi[3])
{
    for (int k = 0; k < 3; ++k) {
        lo[k] = INFINITY;
        hi[k] = -INFINITY;
    }
    for (std::uint32_t i = 0; i < vertexCount; ++i)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], positions[4 * static_cast<std::size_t>(i) + k]);
            hi[k] = std::max(hi[k], positions[4 * static_cast<std::size_t>(i) + k]);
        }
}

Bvh4 BuildBvh4LbvhSerial(const float* positions, const std::uint32_t* indices, std::uint32_t n,
                         std::uint32_t targetClusters, LbvhBuildInfo* info)
{
    if (n < 2 || n > (1U << 28)) throw std::invalid_argument("LBVH needs 2 .. 2^28 triangles");
    std::uint32_t maxIndex = 0;
    for (std::size_t i = 0; i < 3 * static_cast<std::size_t>(n); ++i) maxIndex = std::max(maxIndex, indices[i]);
    float lo[3], hi[3];
    VertexBounds(positions, maxIndex + 1, lo, hi);
    float scale[3];
    lbvh::MortonScale(lo, hi, scale);

    std::vector<std::uint32_t> morton(n), ids(n);
    std::vector<Box4> triBox(n);
    for (std::uint32_t i = 0; i < n; ++i) {
        morton[i] = lbvh::PrepareTriangle(positions, indices, i, lo, scale, triBox[i]);
        ids[i] = i;
    }
    // The device sort is a stable LSD radix sort: equal codes keep id order.
    std::stable_sort(ids.begin(), ids.end(), [&](std::uint32_t a, std::uint32_t b) { return morton[a] < morton[b]; });
    std::vector<std::uint32_t> keys(n);
    for (std::uint32_t k = 0; k < n; ++k) keys[k] = morton[ids[k]];

    const std::uint32_t S = FrontierThreshold(n, targetClusters);
    const std::uint32_t cap = FrontierCapacity(n, S);
    const std::size_t inner = n - 1 + static_cast<std::size_t>(cap);
    std::vector<std::uint32_t> child2(2 * inner, lbvh::kNone), range2(2 * inner, 0U), parent(n, lbvh::kNone),
        leafParent(n, lbvh::kNone), arrivals(n, 0U);
    std::vector<Box4> nodeBox(inner);
    std::vector<float> nodeCost(n - 1, 0.0F);
    std::vector<std::uint32_t> leafFlag(inner, 0U);  // SAH top nodes (ids >= n - 1) stay 0
    for (std::uint32_t i = 0; i + 1 < n; ++i)
        lbvh::RadixTreeNode(keys.data(), static_cast<int>(n), static_cast<int>(i), child2.data(), parent.data(),
                            leafParent.data(), range2.data());
    parent[0] = lbvh::kNone;
    for (std::uint32_t k = 0; k < n; ++k)
        lbvh::BottomUpLeaf<SerialAtomics>(k, child2.data(), parent.data(), leafParent.data(), arrivals.data(),
                                          ids.data(), triBox.data(), range2.data(), nodeBox.data(), nodeCost.data(),
                                          leafFlag.data());

    std::uint32_t root = 0;
    if (info) *info = LbvhBuildInfo{};
    if (S && n > S) {
        std::vector<FrontierItem> items(cap);
        std::uint32_t count = 0;
        for (const float maxArea : {FrontierMaxArea(lo, hi, n, S), INFINITY}) {
            count = 0;
            for (std::uint32_t i = 0; i + 1 < n; ++i)
                lbvh::FrontierNode<SerialAtomics>(i, S, maxArea, child2.data(), range2.data(), nodeBox.data(),
                                                  triBox.data(), ids.data(), &count, cap, items.data());
            if (count <= cap) break;
        }
        if (count >= 2 && count <= cap) {
            items.resize(count);
            std::sort(items.begin(), items.end(), [](const FrontierItem& a, const FrontierItem& b) { return a.ref < b.ref; });
            std::vector<std::uint32_t> tc, tr;
            std::vector<Box4> tb;
            root = BuildSahTop(items, n - 1, tc, tr, tb);
            std::copy(tc.begin(), tc.end(), child2.begin() + 2 * (static_cast<std::size_t>(n) - 1));
            std::copy(tr.begin(), tr.end(), range2.begin() + 2 * (static_cast<std::size_t>(n) - 1));
            std::copy(tb.begin(), tb.end(), nodeBox.begin() + (static_cast<std::size_t>(n) - 1));
            if (info) info->topClusters = count;
        }
    }

    const std::uint32_t nodeCap = n - 1;
    Bvh4 out;
    out.nodes.resize(nodeCap);
    std::vector<std::uint32_t> source(nodeCap, 0U), nodeStack(nodeCap, 0U);
    std::uint32_t counters[4] = {1U, 0U, 0U, 0U};
    source[0] = root;
    std::vector<std::uint32_t> levels{0U};
    std::uint32_t begin = 0, end = 1, level = 0;
    while (begin < end) {
        for (std::uint32_t id = begin; id < end; ++id)
            lbvh::CollapseNode<SerialAtomics>(id, level, source.data(), child2.data(), range2.data(), nodeBox.data(),
                                              triBox.data(), ids.data(), leafFlag.data(), nodeStack.data(), counters,
                                              nodeCap, out.nodes.data());
        if (counters[3]) throw std::runtime_error("LBVH collapse overflowed its node pool");
        levels.push_back(end);
        begin = end;
        end = counters[0];
        ++level;
    }
    out.nodes.resize(counters[0]);
    out.maxStack = counters[1];
    out.depth = counters[2];
    out.triangles.resize(12 * static_cast<std::size_t>(n));
    for (std::uint32_t k = 0; k < n; ++k) lbvh::WriteSlot(positions, indices, ids[k], out.triangles.data() + 12 * static_cast<std::size_t>(k));
    if (info) info->levelBegin = levels;
    return out;
}

} // namespace hdcx
