// Kaiser-Bessel gridding (reference ReconData/GridLut.h:9-41 and the CUDA
// cuGridLut.h:31-57 it dispatches to).
//
// Same interface as the reference.  plan() builds the LUT, normalises the
// trajectory (ReconData::normalizeTraj) and builds the device gridding plan
// once; execute() grids every coil in one launch and returns an ImageData
// whose channels stay in HBM until read.
#ifndef GRIDLUT_H
#define GRIDLUT_H

#include "ReconData.h"
#include "ImageData.h"
#include "fastrecon_hip.h"

class ConvKernel;

template<typename T>
class GridLut
{
protected:
    GridLut(ReconData<T> &reconData);

public:
    virtual ~GridLut();
    static std::shared_ptr<GridLut<T>> Create(ReconData<T> &reconData);

    virtual void plan(unsigned reconSize, float overGridFactor, float kWidth, unsigned klength = 32);
    virtual std::shared_ptr<ImageData<T>> execute();
    void setNumOfThreads(unsigned threads) {
        m_num_threads = threads;
    }
    void setIndex(int index) {
        m_index = index;
    }
    // cuGridLut::setNumOfPartitions (cuGridLut.h:40-42): accepted for
    // compatibility; the plan is resident in HBM, so there are no partitions.
    void setNumOfPartitions(int partitions) {
        m_partitions = partitions;
    }
    // ---- MI355X backend (not in the reference interface) ----
    void setTune(const fr_tune &tune) {
        m_tune = tune;
        m_has_tune = true;
    }
    const fr_plan *devicePlan() const { return m_plan.get(); }
    unsigned gridSize() const { return m_gridSize; }
    const ConvKernel *kernel() const { return m_kernel; }
    // true when the last plan() reused a device plan built for the same
    // trajectory and parameters (by this or another GridLut)
    bool planShared() const { return m_plan_shared; }
    // Plans built by plan() in this process and still held (by a GridLut or
    // as the most recent one, which the process keeps like the reference's
    // static map, cuGridLut.h:50-56).
    static int sharedPlanCount();
    static void releaseSharedPlans();

protected:
    ComplexVector<T> griddingChannel(int channel);

    ReconData<T> &m_associatedData;
    unsigned m_dim = 0;
    unsigned m_gridSize = 0;
    ConvKernel *m_kernel = nullptr;

    unsigned m_num_threads = 1;
    int m_index = 0;
    int m_partitions = 1;

private:
    std::shared_ptr<fr_plan> m_plan;
    fr_tune m_tune{};
    bool m_has_tune = false;
    bool m_plan_shared = false;
    // last execute() result reused by griddingChannel while the data and
    // the plan are unchanged
    std::shared_ptr<ImageData<T>> m_channel_cache;
    unsigned long long m_cache_version = 0;
    const fr_plan *m_cache_plan = nullptr;
};

#endif // GRIDLUT_H
