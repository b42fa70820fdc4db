// mt_jump.hip -- device copies of the std::mt19937 jump polynomials
// (mt_jump_host.hpp) as batched coefficient lists (mt_dev.hpp, mt_jump), made
// once per device and set of distances and kept for the life of the process.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <thread>
#include <mutex>
#include <utility>
#include <vector>

#include "mt_dev.hpp"
#include "mt_jump_host.hpp"
#include "pgh_internal.hpp"

namespace pgh
{

// The batched index list of z^d mod phi (layout: mt_dev.hpp, kMtIdxStride).
static pgh_status index_list(uint64_t d, std::vector<uint16_t> &out)
{
    const mtj::poly &p = mtj::jump_poly(d);
    out.clear();
    uint32_t first = 0, inb = 0, last = 0;
    for (uint32_t i = 0; i < mtj::kDeg; ++i) {
        if (!mtj::bit(p, i)) continue;
        if (inb > 0u && i - first > dev::kMtIdxSpan) {
            while (inb < 8u) {
                out.push_back(static_cast<uint16_t>(last | dev::kMtIdxSkip));
                ++inb;
            }
        }
        if (inb == 8u) inb = 0u;
        if (inb == 0u) first = i;
        out.push_back(static_cast<uint16_t>(i));
        last = i;
        ++inb;
    }
    while (inb > 0u && inb < 8u) {
        out.push_back(static_cast<uint16_t>(last | dev::kMtIdxSkip));
        ++inb;
    }
    if (out.size() > dev::kMtIdxStride) return fail(PGH_DEVICE_ERROR, "jump polynomial index list too long");
    return PGH_OK;
}

pgh_status jump_table_bytes(const std::vector<uint64_t> &dists, std::vector<uint8_t> &h)
{
    if (!mtj::the_field().ok) return fail(PGH_DEVICE_ERROR, "mt19937 characteristic polynomial not found");
    const size_t head = jump_table_head(dists.size());
    h.assign(head + dists.size() * dev::kMtIdxStride * 2u, 0u);
    mtj::prefetch_jump_polys(dists, std::min(8u, std::thread::hardware_concurrency()));
    std::vector<uint16_t> lst;
    for (size_t k = 0; k < dists.size(); ++k) {
        PGH_TRY(index_list(dists[k], lst));
        const uint32_t n = static_cast<uint32_t>(lst.size());
        std::memcpy(h.data() + 4u * k, &n, 4u);
        std::memcpy(h.data() + head + k * dev::kMtIdxStride * 2u, lst.data(), lst.size() * 2u);
    }
    return PGH_OK;
}

pgh_status jump_polys_multiples(uint64_t S, uint64_t count)
{
    if (!mtj::the_field().ok) return fail(PGH_DEVICE_ERROR, "mt19937 characteristic polynomial not found");
    mtj::prefetch_multiples(S, count);
    return PGH_OK;
}

pgh_status jump_table_device(const std::vector<uint64_t> &dists, const void **out)
{
    static std::mutex mu;
    static std::map<std::pair<int, std::vector<uint64_t>>, void *> cache;
    int dev = 0;
    PGH_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(mu);
    const auto key = std::make_pair(dev, dists);
    auto it = cache.find(key);
    if (it == cache.end()) {
        std::vector<uint8_t> h;
        PGH_TRY(jump_table_bytes(dists, h));
        void *ptr = nullptr;
        PGH_HIP(hipMalloc(&ptr, h.size()));
        PGH_HIP(hipMemcpy(ptr, h.data(), h.size(), hipMemcpyHostToDevice));
        it = cache.emplace(key, ptr).first;
    }
    *out = it->second;
    return PGH_OK;
}

} // namespace pgh
