// GridLut (reference ReconData/GridLut.cpp:25-78, cuGridLut.cu:28-48,
// 151-226) on the MI355X backend.
#include <cmath>
#include <iostream>

#include "ConvKernel.h"
#include "GridLut.h"
#include "hip_backend.h"

using fastrecon::check;
using fastrecon::ctx;

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
    fr_plan *p = nullptr;
    m_plan.reset();
    check(fr_plan_create(ctx(), s, m_gridSize, kWidth, lut->data(), (int)lut->size(),
                         m_has_tune ? &m_tune : nullptr, &p),
          "gridding plan");
    m_plan = std::shared_ptr<fr_plan>(p, fr_plan_destroy);
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

template<typename T>
ComplexVector<T> GridLut<T>::griddingChannel(int channel)
{
    auto img = execute();
    return *img->getChannelImage(channel);
}

template class GridLut<float>;
