// GridLut (reference ReconData/GridLut.cpp:25-78, cuGridLut.cu:28-48,
// 151-226) on the MI355X backend.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <iostream>
#include <mutex>
#include <thread>
#include <vector>

#include "ConvKernel.h"
#include "GridLut.h"
#include "hip_backend.h"

using fastrecon::check;
using fastrecon::ctx;

namespace {

// Device plans are shared between GridLut instances whose trajectory (as
// normalised), grid size, kernel, LUT and tuning are the same, as the
// reference shares its gridding map across instances and frames
// (cuGridLut.cu:156-159, cuGridLut.h:50-56, GridLut.h:24-26).  Unlike the
// reference, which keeps the map of the first trajectory it sees, the key
// includes a 64-bit content hash of the trajectory, so another trajectory
// gets its own plan.
struct PlanKey {
    unsigned long long traj_hash;
    long long samples;
    int dim;
    unsigned grid;
    float kwidth, og;
    unsigned klength;
    bool has_tune;
    fr_tune tune;

    bool operator==(const PlanKey &o) const
    {
        return traj_hash == o.traj_hash && samples == o.samples && dim == o.dim && grid == o.grid &&
               kwidth == o.kwidth && og == o.og && klength == o.klength && has_tune == o.has_tune &&
               (!has_tune || std::memcmp(&tune, &o.tune, sizeof(fr_tune)) == 0);
    }
};

struct PlanRegistry {
    std::mutex mutex;
    std::vector<std::pair<PlanKey, std::weak_ptr<fr_plan>>> plans;
    std::shared_ptr<fr_plan> last;      // the most recent plan stays alive
};

PlanRegistry &registry()
{
    (void)ctx();                        // constructed first, so destroyed after the plans
    static PlanRegistry r;
    return r;
}

// 64-bit hash of the first `dim` coordinates of every point (their float
// bits), in four interleaved streams over the point range.
template<typename T>
unsigned long long traj_hash(const typename ReconData<T>::TrajVector &traj, int dim)
{
    const std::size_t n = traj.size();
    const int parts = (int)std::min<std::size_t>(8, std::max<std::size_t>(1, n / 65536));
    std::vector<unsigned long long> h(parts, 0);
    auto work = [&](int q) {
        const std::size_t a = n * q / parts, b = n * (q + 1) / parts;
        unsigned long long s0 = 0x9e3779b97f4a7c15ull ^ (unsigned long long)q, s1 = 0xc2b2ae3d27d4eb4full;
        for (std::size_t i = a; i < b; ++i) {
            for (int d = 0; d < dim; ++d) {
                uint32_t w;
                std::memcpy(&w, &traj[i].x[d], 4);
                s0 = (s0 ^ w) * 0x100000001b3ull;
                s1 = (s1 + w) * 0xff51afd7ed558ccdull;
                s1 ^= s1 >> 29;
            }
        }
        h[q] = s0 ^ (s1 * 0x9e3779b97f4a7c15ull) ^ (unsigned long long)(b - a);
    };
    std::vector<std::thread> th;
    for (int q = 1; q < parts; ++q) th.emplace_back(work, q);
    work(0);
    for (auto &t : th) t.join();
    unsigned long long r = 0xcbf29ce484222325ull;
    for (unsigned long long v : h) r = (r ^ v) * 0x100000001b3ull;
    return r;
}

}  // namespace

template<typename T>
GridLut<T>::GridLut(ReconData<T> &reconData)
    : m_associatedData(reconData)
{
}

template<typename T>
GridLut<T>::~GridLut()
{
    if (m_kernel != nullptr)
        delete m_kernel;
}

// GridLut.cpp:25-39: the factory; this library has one (device) backend.
template<typename T>
std::shared_ptr<GridLut<T>> GridLut<T>::Create(ReconData<T> &reconData)
{
    return std::shared_ptr<GridLut<T>>(new GridLut<T>(reconData));
}

// GridLut.cpp:41-51 (+ cuGridLut.cu:151-226): grid size, LUT, trajectory
// normalisation, then the device plan for this trajectory.
template<typename T>
void GridLut<T>::plan(unsigned reconSize, float overGridFactor, float kWidth, unsigned klength)
{
    m_dim = m_associatedData.rcDim();
    m_gridSize = unsigned(reconSize * overGridFactor);
    if (m_kernel != nullptr)
        delete m_kernel;
    m_kernel = new ConvKernel(kWidth, overGridFactor, klength);

    m_associatedData.normalizeTraj(m_gridSize);

    fr_samples *s = m_associatedData.deviceSamples();
    const std::vector<float> *lut = m_kernel->getKernelData();
    PlanKey key{traj_hash<T>(m_associatedData.getTraj(), (int)m_dim), (long long)m_associatedData.dataSize(),
                (int)m_dim, m_gridSize, kWidth, overGridFactor, klength, m_has_tune, m_has_tune ? m_tune : fr_tune{}};
    m_plan.reset();
    m_plan_shared = false;
    PlanRegistry &reg = registry();
    std::lock_guard<std::mutex> lock(reg.mutex);             // plan() is serialised (cuGridLut.cu:156)
    for (auto it = reg.plans.begin(); it != reg.plans.end();) {
        if (it->second.expired()) {
            it = reg.plans.erase(it);
            continue;
        }
        if (it->first == key) {
            m_plan = it->second.lock();
            m_plan_shared = true;
            break;
        }
        ++it;
    }
    if (!m_plan) {
        fr_plan *p = nullptr;
        check(fr_plan_create(ctx(), s, m_gridSize, kWidth, lut->data(), (int)lut->size(),
                             m_has_tune ? &m_tune : nullptr, &p),
              "gridding plan");
        m_plan = std::shared_ptr<fr_plan>(p, fr_plan_destroy);
        reg.plans.emplace_back(key, m_plan);
    }
    reg.last = m_plan;
}

template<typename T>
int GridLut<T>::sharedPlanCount()
{
    PlanRegistry &reg = registry();
    std::lock_guard<std::mutex> lock(reg.mutex);
    int n = 0;
    for (auto &e : reg.plans) n += e.second.expired() ? 0 : 1;
    return n;
}

template<typename T>
void GridLut<T>::releaseSharedPlans()
{
    PlanRegistry &reg = registry();
    std::lock_guard<std::mutex> lock(reg.mutex);
    reg.last.reset();
}

// GridLut.cpp:53-78 / cuGridLut.cu:28-48: all coils in one launch; the
// returned image keeps its channels in HBM until they are read.
template<typename T>
std::shared_ptr<ImageData<T>> GridLut<T>::execute()
{
    if (!m_plan) {
        std::cerr << "Error: run GridLut<T>::plan first." << std::endl;   // cuGridLut.cu:31-34
        exit(1);
    }
    fr_samples *s = m_associatedData.deviceSamples();
    const int C = m_associatedData.channels();
    fr_grids *g = nullptr;
    check(fr_grids_create(ctx(), (int)m_dim, m_gridSize, C, &g), "allocate grids");
    fr_status st = fr_grid_execute(m_plan.get(), s, g);
    if (st != FR_OK) {
        fr_grids_destroy(g);
        check(st, "gridding");
    }
    auto img = new ImageData<T>(m_dim, {m_gridSize, m_gridSize, m_gridSize});
    img->adoptDeviceGrids(g, C);
    return std::shared_ptr<ImageData<T>>(img);
}

// One channel's grid (GridLut.cpp:80-147).  All coils grid in one launch,
// so the pass is kept and later channels are read from it while the data
// and the plan are unchanged: a loop over channels costs one pass.
template<typename T>
ComplexVector<T> GridLut<T>::griddingChannel(int channel)
{
    if (!m_channel_cache || m_cache_version != m_associatedData.version() || m_cache_plan != m_plan.get()) {
        m_channel_cache.reset();
        m_channel_cache = execute();
        m_cache_version = m_associatedData.version();
        m_cache_plan = m_plan.get();
    }
    return *m_channel_cache->getChannelImage(channel);
}

template class GridLut<float>;
