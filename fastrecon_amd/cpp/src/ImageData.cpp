// ImageData (reference ImageData/ImageData.cpp:6-111) with device residency.
#include <iostream>

#include "ImageData.h"
#include "hip_backend.h"

using fastrecon::check;
using fastrecon::ctx;

template<typename T>
ImageData<T>::ImageData(const int dim, const ImageSize &size)
    : m_dim(dim), m_size(size)
{
    if (dim == 2)
        m_size.z = 1;          // ImageData.cpp:10-11
}

template<typename T>
ImageData<T>::ImageData(const ImageData<T> &imageData)
{
    copy(imageData);
}

template<typename T>
ImageData<T>::ImageData(ImageData<T> &&imageData)
{
    move(imageData);
}

template<typename T>
ImageData<T>::~ImageData() {}

template<typename T>
ImageData<T> &ImageData<T>::operator=(const ImageData<T> &imageData)
{
    copy(imageData);
    return *this;
}

template<typename T>
ImageData<T> &ImageData<T>::operator=(ImageData<T> &&imageData)
{
    move(imageData);
    return *this;
}

template<typename T>
std::size_t ImageData<T>::dataSize() const
{
    if (m_dim == 3)
        return (std::size_t)m_size.x * m_size.y * m_size.z;
    else
        return (std::size_t)m_size.x * m_size.y;
}

template<typename T>
void ImageData<T>::addChannelImage(ComplexVector<T> &&image)
{
    if (image.size() != dataSize())              // ImageData.cpp:56-60
    {
        std::cerr << "Error: ImageData wrong size!" << std::endl;
        exit(1);
    }
    syncToHost();
    dropDevice();
    m_data_multichannel.push_back(std::move(image));
    m_channels = m_data_multichannel.size();
}

template<typename T>
const ComplexVector<T> *ImageData<T>::getChannelImage(int channel) const
{
    if (channel < channels()) {
        syncToHost();
        return &m_data_multichannel[channel];
    }
    return nullptr;
}

template<typename T>
ComplexVector<T> *ImageData<T>::getChannelImage(int channel)
{
    if (channel < channels()) {
        syncToHost();
        m_dev_valid = false;       // the caller may write through the pointer
        return &m_data_multichannel[channel];
    }
    return nullptr;
}

template<typename T>
void ImageData<T>::materializeFft()
{
    if (m_fft_pending && m_dev && m_dev_valid)
        check(fr_fft_execute(m_fft_pending.get(), m_dev.get()), "FFT");
    m_fft_pending.reset();
}

template<typename T>
void ImageData<T>::materializeShift()
{
    materializeFft();
    if (m_shift_pending && m_dev && m_dev_valid) {
        check(fr_fftshift(ctx(), m_dev.get()), "fftShift");
        m_shift_pending = false;
    }
}

template<typename T>
void ImageData<T>::syncToHost() const
{
    if (m_host_valid) return;
    auto *self = const_cast<ImageData<T> *>(this);
    self->materializeShift();
    m_data_multichannel.resize(m_channels);
    for (int c = 0; c < m_channels; ++c) {
        m_data_multichannel[c].resize(dataSize());
        check(fr_grids_download(m_dev.get(), c, m_data_multichannel[c].data()), "download channel");
    }
    m_host_valid = true;
}

template<typename T>
void ImageData<T>::dropDevice()
{
    m_dev.reset();
    m_dev_valid = false;
    m_shift_pending = false;
    m_fft_pending.reset();
}

template<typename T>
fr_grids *ImageData<T>::deviceGrids()
{
    if (!(m_dev && m_dev_valid)) {
        if (m_size.x != m_size.y || (m_dim == 3 && m_size.z != m_size.x)) {
            std::cerr << "Error: device images must be square/cubic" << std::endl;
            exit(1);
        }
        syncToHost();
        fr_grids *g = nullptr;
        check(fr_grids_create(ctx(), m_dim, m_size.x, m_channels, &g), "allocate device grids");
        m_dev = std::shared_ptr<fr_grids>(g, fr_grids_destroy);
        for (int c = 0; c < m_channels; ++c)
            check(fr_grids_upload(g, c, m_data_multichannel[c].data()), "upload channel");
        m_dev_valid = true;
        m_shift_pending = false;
        m_fft_pending.reset();
    }
    m_host_valid = false;
    return m_dev.get();
}

template<typename T>
const fr_grids *ImageData<T>::deviceGrids() const
{
    // read-only use: keep the host copy valid if it was
    auto *self = const_cast<ImageData<T> *>(this);
    const bool host = m_host_valid;
    fr_grids *g = self->deviceGrids();
    m_host_valid = host;
    return g;
}

template<typename T>
void ImageData<T>::adoptDeviceGrids(fr_grids *grids, int channels)
{
    m_dev = std::shared_ptr<fr_grids>(grids, fr_grids_destroy);
    m_dev_valid = true;
    m_shift_pending = false;
    m_fft_pending.reset();
    m_channels = channels;
    m_data_multichannel.clear();
    m_host_valid = false;
}

template<typename T>
void ImageData<T>::copy(const ImageData<T> &imageData)
{
    m_dim = imageData.dim();
    m_size = imageData.imageSize();
    m_channels = imageData.m_channels;
    if (imageData.m_dev && imageData.m_dev_valid && !imageData.m_host_valid) {
        // device-resident source: copy in HBM
        fr_grids *g = nullptr;
        const fr_grids *src = imageData.m_dev.get();
        check(fr_grids_create(ctx(), m_dim, m_size.x, m_channels, &g), "allocate device grids");
        check(fr_grids_copy(ctx(), src, g), "copy device grids");
        m_dev = std::shared_ptr<fr_grids>(g, fr_grids_destroy);
        m_dev_valid = true;
        m_shift_pending = imageData.m_shift_pending;
        m_fft_pending = imageData.m_fft_pending;
        m_data_multichannel.clear();
        m_host_valid = false;
    } else {
        imageData.syncToHost();
        dropDevice();
        m_data_multichannel = imageData.m_data_multichannel;
        m_host_valid = true;
    }
    std::cout << "-- ImageData: copy --" << std::endl;      // ImageData.cpp:94
}

template<typename T>
void ImageData<T>::move(ImageData<T> &imageData)
{
    m_dim = imageData.m_dim;
    m_size = imageData.m_size;
    m_channels = imageData.m_channels;
    m_data_multichannel = std::move(imageData.m_data_multichannel);
    m_dev = std::move(imageData.m_dev);
    m_host_valid = imageData.m_host_valid;
    m_dev_valid = imageData.m_dev_valid;
    m_shift_pending = imageData.m_shift_pending;
    m_fft_pending = std::move(imageData.m_fft_pending);
    imageData.m_data_multichannel.clear();
    imageData.m_dev.reset();
    imageData.m_dim = 0;
    imageData.m_size = {0, 0, 0};
    imageData.m_channels = 0;
    imageData.m_host_valid = true;
    imageData.m_dev_valid = false;
    imageData.m_shift_pending = false;
    imageData.m_fft_pending.reset();
    std::cout << "-- ImageData: move --" << std::endl;      // ImageData.cpp:110
}

template class ImageData<float>;
