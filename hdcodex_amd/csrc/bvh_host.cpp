// Host BVH4 builder: binned SAH on a BVH2, then collapse to four-wide nodes.
// Used for small scenes and as the reference layout the device LBVH builder
// is checked against (both emit the same node and triangle formats).
#include "bvh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <stdexcept>
#include <string>

namespace hdcx {
namespace {

struct Box {
    float lo[3] = {INFINITY, INFINITY, INFINITY};
    float hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    void Grow(const Box& b)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], b.lo[k]);
            hi[k] = std::max(hi[k], b.hi[k]);
        }
    }
    void GrowPoint(const float* p)
    {
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], p[k]);
            hi[k] = std::max(hi[k], p[k]);
        }
    }
    [[nodiscard]] double Area() const
    {
        if (!(lo[0] <= hi[0])) return 0.0;
        const double dx = double(hi[0]) - lo[0], dy = double(hi[1]) - lo[1], dz = double(hi[2]) - lo[2];
        return 2.0 * (dx * dy + dy * dz + dz * dx);
    }
};

struct Ref {
    Box box;
    float c[3];
    std::uint32_t id;
};

struct Node2 {
    Box box;
    std::uint32_t left = 0, right = 0;   // inner children
    std::uint32_t first = 0, count = 0;  // leaf range in refs (count > 0)
};

constexpr int kBins = 16;
constexpr std::uint32_t kLeafSize = 4;

class Builder {
public:
    Builder(const float* pos, const std::uint32_t* idx, std::uint32_t n) : positions(pos), indices(idx), count(n) {}

    Bvh4 Build()
    {
        refs.resize(count);
        for (std::uint32_t t = 0; t < count; ++t) {
            Ref& r = refs[t];
            r.id = t;
            for (int c = 0; c < 3; ++c) r.box.GrowPoint(&positions[static_cast<std::size_t>(indices[t * 3 + c]) * 4]);
            for (int k = 0; k < 3; ++k) r.c[k] = 0.5F * (r.box.lo[k] + r.box.hi[k]);
        }
        nodes.reserve(count > 0 ? 2 * count : 1);
        nodes.emplace_back();
        struct Task { std::uint32_t node, begin, end; };
        std::vector<Task> tasks;
        tasks.push_back({0, 0, count});
        while (!tasks.empty()) {
            const Task task = tasks.back();
            tasks.pop_back();
            Box box, cbox;
            for (std::uint32_t i = task.begin; i < task.end; ++i) {
                box.Grow(refs[i].box);
                cbox.GrowPoint(refs[i].c);
            }
            nodes[task.node].box = box;
            const std::uint32_t n = task.end - task.begin;
            std::uint32_t mid = task.begin;
            if (n <= 2 || !Split(task.begin, task.end, box, cbox, mid)) {
                if (n <= kLeafSize) {
                    nodes[task.node].first = task.begin;
                    nodes[task.node].count = n;
                    continue;
                }
                // Degenerate centroids: split by position in the range.
                mid = task.begin + n / 2;
            }
            const std::uint32_t l = static_cast<std::uint32_t>(nodes.size());
            nodes.emplace_back();
            nodes.emplace_back();
            nodes[task.node].left = l;
            nodes[task.node].right = l + 1;
            tasks.push_back({l + 1, mid, task.end});
            tasks.push_back({l, task.begin, mid});
        }
        Bvh4 out;
        if (count == 0) return out;
        out.nodes.emplace_back();
        if (nodes[0].count > 0) {
            // A single leaf becomes the only child of an inner root: the
            // traversal always starts at an inner node.
            Bvh4Node root = EmptyNode();
            SetChild(root, 0, nodes[0], out);
            out.nodes[0] = root;
            MeasureBvh4(out);
            return out;
        }
        Collapse(0, 0, out);
        MeasureBvh4(out);
        return out;
    }

private:
    // Binned SAH split; returns false when a leaf is cheaper (and fits) or
    // no split plane separates the centroids.
    bool Split(std::uint32_t begin, std::uint32_t end, const Box& box, const Box& cbox, std::uint32_t& mid)
    {
        const std::uint32_t n = end - begin;
        double bestCost = std::numeric_limits<double>::infinity();
        int bestAxis = -1, bestBin = -1;
        for (int axis = 0; axis < 3; ++axis) {
            const float extent = cbox.hi[axis] - cbox.lo[axis];
            if (!(extent > 0.0F)) continue;
            Box bins[kBins];
            std::uint32_t counts[kBins] = {};
            const float scale = kBins / extent;
            for (std::uint32_t i = begin; i < end; ++i) {
                int b = static_cast<int>((refs[i].c[axis] - cbox.lo[axis]) * scale);
                b = std::clamp(b, 0, kBins - 1);
                bins[b].Grow(refs[i].box);
                ++counts[b];
            }
            double rightArea[kBins];
            std::uint32_t rightCount[kBins];
            Box acc;
            std::uint32_t cnt = 0;
            for (int b = kBins - 1; b > 0; --b) {
                acc.Grow(bins[b]);
                cnt += counts[b];
                rightArea[b] = acc.Area();
                rightCount[b] = cnt;
            }
            acc = Box();
            cnt = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                acc.Grow(bins[b]);
                cnt += counts[b];
                if (cnt == 0 || rightCount[b + 1] == 0) continue;
                const double cost = acc.Area() * cnt + rightArea[b + 1] * rightCount[b + 1];
                if (cost < bestCost) {
                    bestCost = cost;
                    bestAxis = axis;
                    bestBin = b;
                }
            }
        }
        if (bestAxis < 0) return false;
        const double leafCost = box.Area() * n;
        if (n <= kLeafSize && bestCost * 1.0 + box.Area() * 1.2 >= leafCost) return false;
        const float extent = cbox.hi[bestAxis] - cbox.lo[bestAxis];
        const float scale = kBins / extent;
        auto it = std::stable_partition(refs.begin() + begin, refs.begin() + end, [&](const Ref& r) {
            int b = static_cast<int>((r.c[bestAxis] - cbox.lo[bestAxis]) * scale);
            b = std::clamp(b, 0, kBins - 1);
            return b <= bestBin;
        });
        mid = static_cast<std::uint32_t>(it - refs.begin());
        return mid > begin && mid < end;
    }

    static void PadBox(Box& b)
    {
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = std::nextafter(std::nextafter(b.lo[k], -INFINITY), -INFINITY);
            b.hi[k] = std::nextafter(std::nextafter(b.hi[k], INFINITY), INFINITY);
        }
    }

    static Bvh4Node EmptyNode()
    {
        Bvh4Node node{};
        for (int i = 0; i < 4; ++i) {
            node.lox[i] = node.loy[i] = node.loz[i] = INFINITY;
            node.hix[i] = node.hiy[i] = node.hiz[i] = -INFINITY;
            node.child[i] = kEmptyChild;
        }
        return node;
    }

    // Fills child slot i of `node` from BVH2 node c: a leaf writes its
    // triangles to the slot array; an inner child reserves a BVH4 node and
    // returns true (to be collapsed by the caller).
    bool SetChild(Bvh4Node& node, int i, const Node2& c, Bvh4& out)
    {
        Box b = c.box;
        PadBox(b);
        node.lox[i] = b.lo[0];
        node.loy[i] = b.lo[1];
        node.loz[i] = b.lo[2];
        node.hix[i] = b.hi[0];
        node.hiy[i] = b.hi[1];
        node.hiz[i] = b.hi[2];
        if (c.count > 0) {
            const std::uint32_t slot = static_cast<std::uint32_t>(out.triangles.size() / 12);
            for (std::uint32_t j = 0; j < c.count; ++j) {
                const std::uint32_t tri = refs[c.first + j].id;
                for (int v = 0; v < 3; ++v) {
                    const float* p = &positions[static_cast<std::size_t>(indices[tri * 3 + v]) * 4];
                    float w = 0.0F;
                    if (v == 0) std::memcpy(&w, &tri, 4);
                    out.triangles.insert(out.triangles.end(), {p[0], p[1], p[2], w});
                }
            }
            node.child[i] = EncodeLeaf(slot, c.count);
            return false;
        }
        node.child[i] = static_cast<std::uint32_t>(out.nodes.size());
        out.nodes.emplace_back();
        return true;
    }

    void Collapse(std::uint32_t n2, std::uint32_t n4, Bvh4& out)
    {
        std::uint32_t kids[4];
        int k = 0;
        kids[k++] = nodes[n2].left;
        kids[k++] = nodes[n2].right;
        while (k < 4) {
            int best = -1;
            double bestArea = -1.0;
            for (int i = 0; i < k; ++i) {
                const Node2& c = nodes[kids[i]];
                if (c.count == 0 && c.box.Area() > bestArea) {
                    bestArea = c.box.Area();
                    best = i;
                }
            }
            if (best < 0) break;
            const std::uint32_t inner = kids[best];
            kids[best] = nodes[inner].left;
            kids[k++] = nodes[inner].right;
        }
        Bvh4Node node = EmptyNode();
        std::uint32_t pending[4];
        int pendingCount = 0;
        for (int i = 0; i < k; ++i) {
            if (SetChild(node, i, nodes[kids[i]], out)) pending[pendingCount++] = static_cast<std::uint32_t>(i);
        }
        out.nodes[n4] = node;
        // Recurse after the node is stored (out.nodes may reallocate).
        std::uint32_t childNode[4], childN2[4];
        for (int p = 0; p < pendingCount; ++p) {
            childNode[p] = node.child[pending[p]];
            childN2[p] = kids[pending[p]];
        }
        for (int p = 0; p < pendingCount; ++p) Collapse(childN2[p], childNode[p], out);
    }

    const float* positions;
    const std::uint32_t* indices;
    std::uint32_t count;
    std::vector<Ref> refs;
    std::vector<Node2> nodes;
};

} // namespace

Bvh4 BuildBvh4Host(const float* positions, const std::uint32_t* indices, std::uint32_t triangleCount)
{
    Builder b(positions, indices, triangleCount);
    return b.Build();
}

void MeasureBvh4(Bvh4& bvh)
{
    bvh.maxStack = 0;
    bvh.depth = 0;
    if (bvh.nodes.empty()) return;
    struct Item { std::uint32_t node, stack, depth; };
    std::vector<Item> work{{0, 0, 1}};
    while (!work.empty()) {
        const Item it = work.back();
        work.pop_back();
        const Bvh4Node& n = bvh.nodes[it.node];
        std::uint32_t valid = 0;
        for (std::uint32_t c : n.child)
            if (c != kEmptyChild) ++valid;
        const std::uint32_t stack = it.stack + (valid > 0 ? valid - 1 : 0);
        bvh.maxStack = std::max(bvh.maxStack, stack);
        bvh.depth = std::max(bvh.depth, it.depth);
        for (std::uint32_t c : n.child)
            if (c != kEmptyChild && (c & kLeafFlag) == 0) work.push_back({c, stack, it.depth + 1});
    }
}

void ValidateBvh4(const Bvh4& bvh, std::uint32_t triangleCount)
{
    auto fail = [](const std::string& m) { throw std::runtime_error("invalid BVH: " + m); };
    if (triangleCount == 0) {
        if (!bvh.nodes.empty()) fail("nodes without triangles");
        return;
    }
    if (bvh.nodes.empty()) fail("no root");
    const std::size_t slots = bvh.triangles.size() / 12;
    if (bvh.triangles.size() % 12 != 0) fail("slot array not a multiple of 3 float4");
    std::vector<std::uint8_t> seen(triangleCount, 0);
    std::vector<std::uint8_t> visited(bvh.nodes.size(), 0);
    std::vector<std::uint32_t> work{0};
    visited[0] = 1;
    while (!work.empty()) {
        const Bvh4Node& n = bvh.nodes[work.back()];
        work.pop_back();
        for (int i = 0; i < 4; ++i) {
            const std::uint32_t c = n.child[i];
            if (c == kEmptyChild) {
                if (!(n.lox[i] > n.hix[i])) fail("empty slot with a non-inverted box");
                continue;
            }
            if (c & kLeafFlag) {
                const std::uint32_t first = (c >> 3) & 0x0fffffffU, count = (c & 7U) + 1U;
                if (static_cast<std::size_t>(first) + count > slots) fail("leaf beyond the slot array");
                for (std::uint32_t j = 0; j < count; ++j) {
                    std::uint32_t prim;
                    std::memcpy(&prim, &bvh.triangles[(static_cast<std::size_t>(first) + j) * 12 + 3], 4);
                    if (prim >= triangleCount) fail("slot primitive id out of range");
                    if (seen[prim]++) fail("triangle referenced twice");
                }
            } else {
                if (c >= bvh.nodes.size()) fail("inner reference out of range");
                if (visited[c]++) fail("node reached twice");
                work.push_back(c);
            }
        }
    }
    for (std::uint32_t t = 0; t < triangleCount; ++t)
        if (!seen[t]) fail("triangle " + std::to_string(t) + " not in any leaf");
}

} // namespace hdcx
