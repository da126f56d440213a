// ConvKernel (reference ReconData/ConvKernel.cpp:7-46); the table itself is
// computed by fr_kb_lut with the reference's float/double arithmetic.
#include <iostream>

#include "ConvKernel.h"
#include "fastrecon_hip.h"

ConvKernel::ConvKernel(float kWidth, float overGridFactor, int length)
    : m_kWidth(kWidth), m_ogFactor(overGridFactor), m_length(length)
{
    auto *data = new std::vector<float>(m_length);
    if (fr_kb_lut(m_kWidth, m_ogFactor, m_length, data->data(), nullptr) != FR_OK) {
        std::cerr << "Error: invalid kernel (width " << kWidth << ", oversampling " << overGridFactor
                  << ", length " << length << ")" << std::endl;
        exit(1);
    }
    m_kernelData.reset(data);
}

ConvKernel::~ConvKernel() {}

const std::vector<float> *ConvKernel::getKernelData() const
{
    return m_kernelData.get();
}

float ConvKernel::getKernelWidth() const
{
    return m_kWidth;
}
