// Checks of the C++ API (GridLut / ReconData) on the device, run by
// tests/test_gpu_cpp_api.py.  Prints one JSON object.
//
//   plan sharing: two ReconData objects with the same trajectory and
//   different coil data; the second GridLut::plan() reuses the first one's
//   device plan (GridLut.h:24-26, cuGridLut.h:50-56 share one map between
//   instances) and grids bit-identically to a plan built afresh;
//   a third ReconData with another trajectory gets its own plan;
//   griddingChannel over all channels: one gridding pass, channels equal to
//   execute()'s (GridLut.cpp:53-78).
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "GridLut.h"
#include "ReconData.h"

namespace {

// kooshball of `spokes` Fibonacci directions x `readout` samples, DCF |k|^2
std::vector<float> kooshball(int spokes, int readout, float rot)
{
    std::vector<float> t;
    t.reserve((size_t)spokes * readout * 4);
    const double ga = M_PI * (3.0 - std::sqrt(5.0));
    for (int i = 0; i < spokes; ++i) {
        const double z = 1.0 - (2.0 * i + 1.0) / spokes;
        const double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        const double ph = i * ga + rot;
        const double u[3] = {r * std::cos(ph), r * std::sin(ph), z};
        for (int j = 0; j < readout; ++j) {
            const double k = (j - readout / 2) / (double)readout;
            double k2 = 0;
            for (int d = 0; d < 3; ++d) {
                t.push_back((float)(k * u[d]));
                k2 += k * u[d] * k * u[d];
            }
            t.push_back((float)k2);                 // interleaved x, y, z, dcf (ReconData.cpp:146-177)
        }
    }
    return t;
}

ComplexVector<float> coil(size_t n, uint64_t seed)
{
    ComplexVector<float> v(n);
    uint64_t x = seed * 0x9e3779b97f4a7c15ull + 1;
    for (auto &c : v) {
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        const float a = (float)((x >> 11) * (1.0 / 9007199254740992.0)) - 0.5f;
        x ^= x << 13; x ^= x >> 7; x ^= x << 17;
        const float b = (float)((x >> 11) * (1.0 / 9007199254740992.0)) - 0.5f;
        c = {a, b};
    }
    return v;
}

std::shared_ptr<ReconData<float>> make(int spokes, int readout, int C, float rot, uint64_t seed)
{
    auto rd = ReconData<float>::Create(readout, spokes, true);
    rd->loadTraj(kooshball(spokes, readout, rot), 3);
    for (int c = 0; c < C; ++c) rd->addChannelData(coil((size_t)spokes * readout, seed * 100 + c));
    return rd;
}

class CountingGridLut : public GridLut<float>
{
public:
    explicit CountingGridLut(ReconData<float> &rd) : GridLut<float>(rd) {}
    std::shared_ptr<ImageData<float>> execute() override
    {
        ++passes;
        return GridLut<float>::execute();
    }
    ComplexVector<float> channel(int c) { return griddingChannel(c); }
    int passes = 0;
};

bool same(const ComplexVector<float> &a, const ComplexVector<float> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(a[0])) == 0;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char **argv)
{
    const int spokes = argc > 1 ? std::atoi(argv[1]) : 4000;
    const int readout = argc > 2 ? std::atoi(argv[2]) : 128;
    const int C = argc > 3 ? std::atoi(argv[3]) : 4;
    const unsigned image = argc > 4 ? (unsigned)std::atoi(argv[4]) : 64;

    auto a = make(spokes, readout, C, 0.f, 1);
    auto b = make(spokes, readout, C, 0.f, 2);     // same trajectory, other data
    auto d = make(spokes, readout, C, 0.3f, 3);    // another trajectory

    auto ga = GridLut<float>::Create(*a);
    auto t0 = std::chrono::steady_clock::now();
    ga->plan(image, 2.0f, 4.0f, 512);
    const double plan_a = ms_since(t0);
    auto gb = GridLut<float>::Create(*b);
    t0 = std::chrono::steady_clock::now();
    gb->plan(image, 2.0f, 4.0f, 512);
    const double plan_b = ms_since(t0);
    const bool shared_b = gb->planShared() && gb->devicePlan() == ga->devicePlan();
    auto gd = GridLut<float>::Create(*d);
    gd->plan(image, 2.0f, 4.0f, 512);
    const bool own_d = !gd->planShared() && gd->devicePlan() != ga->devicePlan();
    const int count3 = GridLut<float>::sharedPlanCount();

    auto img_b = gb->execute();
    // b again on a fresh plan: drop every holder of the shared one first
    std::vector<ComplexVector<float>> shared_grid;
    for (int c = 0; c < C; ++c) shared_grid.push_back(*img_b->getChannelImage(c));
    img_b.reset();
    ga.reset();
    gb.reset();
    GridLut<float>::releaseSharedPlans();
    auto gb2 = GridLut<float>::Create(*b);
    gb2->plan(image, 2.0f, 4.0f, 512);
    const bool rebuilt = !gb2->planShared();
    auto img_b2 = gb2->execute();
    bool bitwise = true;
    for (int c = 0; c < C; ++c) bitwise = bitwise && same(shared_grid[c], *img_b2->getChannelImage(c));

    // griddingChannel: one pass for all channels, then a new pass after an update
    CountingGridLut cg(*d);
    cg.plan(image, 2.0f, 4.0f, 512);
    const bool cg_shared = cg.planShared();
    auto full = gd->execute();
    bool chan_equal = true;
    for (int c = 0; c < C; ++c) chan_equal = chan_equal && same(cg.channel(c), *full->getChannelImage(c));
    const int passes_all = cg.passes;
    auto acq = coil(readout, 77);
    d->updateSingleAcquisition(acq.data(), 3, 1);
    auto full2 = gd->execute();
    const bool updated = same(cg.channel(1), *full2->getChannelImage(1)) &&
                         !same(*full->getChannelImage(1), *full2->getChannelImage(1));
    const int passes_after = cg.passes;

    std::printf("{\"plan_first_ms\": %.3f, \"plan_shared_ms\": %.3f, \"shared\": %s, \"other_traj_own_plan\": %s, "
                "\"plans_held\": %d, \"rebuilt_after_release\": %s, \"bitwise_shared_vs_fresh\": %s, "
                "\"channel_loop_shares_plan\": %s, \"channel_loop_passes\": %d, \"channels_equal_execute\": %s, "
                "\"passes_after_update\": %d, \"update_seen\": %s}\n",
                plan_a, plan_b, shared_b ? "true" : "false", own_d ? "true" : "false", count3,
                rebuilt ? "true" : "false", bitwise ? "true" : "false", cg_shared ? "true" : "false", passes_all,
                chan_equal ? "true" : "false", passes_after, updated ? "true" : "false");
    return 0;
}
