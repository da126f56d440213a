// ImageFilter (reference ImageFilter/ImageFilter.cpp:10-282, cuFFT.cpp,
// cuImageFilter.cpp) on the MI355X backend: every operation runs on the
// ImageData's device grids.
#include <iostream>

#include "ImageFilter.h"
#include "hip_backend.h"

using fastrecon::check;
using fastrecon::ctx;

template<typename T>
std::shared_ptr<ImageFilter<T>> ImageFilter<T>::Create(ImageData<T> &imageData)
{
    return std::shared_ptr<ImageFilter<T>>(new ImageFilter<T>(imageData));
}

template<typename T>
ImageFilter<T>::~ImageFilter() = default;

template<typename T>
ImageFilter<T>::ImageFilter(ImageData<T> &imageData)
    : m_associatedData(imageData)
{
}

// ImageFilter.cpp:40-53: an N/2 roll of every axis (even N).  Recorded as
// pending and folded into the next reader of the device data.
template<typename T>
void ImageFilter<T>::fftShift()
{
    auto &d = m_associatedData;
    if (d.imageSize().x % 2 || d.imageSize().y % 2 || (d.dim() == 3 && d.imageSize().z % 2)) {
        std::cerr << "Error: fftShift needs even image dimensions" << std::endl;
        exit(1);
    }
    d.deviceGrids();
    d.setShiftPending(!d.shiftPending());
}

template<typename T>
void ImageFilter<T>::lowFilter(int res)
{
    fr_grids *g = m_associatedData.deviceGrids();
    m_associatedData.materializeShift();
    check(fr_lowfilter(ctx(), g, res), "lowFilter");
}

template<typename T>
void ImageFilter<T>::normalize()
{
    fr_grids *g = m_associatedData.deviceGrids();
    m_associatedData.materializeShift();
    check(fr_normalize(ctx(), g), "normalize");
}

template<typename T>
void ImageFilter<T>::crop(const ImageSize &imageSize)
{
    auto &d = m_associatedData;
    const fr_grids *g = d.deviceGrids();
    d.materializeShift();
    ImageData<T> img(d.dim(), imageSize);
    const std::size_t n = img.dataSize();
    const int C = d.channels();
    void *buf = nullptr;
    check(fr_malloc(ctx(), sizeof(std::complex<T>) * n * C, &buf), "allocate crop");
    check(fr_crop(ctx(), g, imageSize.x, imageSize.y, d.dim() == 3 ? imageSize.z : 1, buf), "crop");
    for (int c = 0; c < C; ++c) {
        ComplexVector<T> out(n);
        check(fr_memcpy_d2h(ctx(), out.data(), static_cast<char *>(buf) + sizeof(std::complex<T>) * n * c,
                            sizeof(std::complex<T>) * n), "download crop");
        img.addChannelImage(std::move(out));
    }
    fr_free(ctx(), buf);
    m_associatedData = std::move(img);
}

template<typename T>
void ImageFilter<T>::SOS(ImageSize reconSize)
{
    SOS(reconSize, nullptr);
}

// ImageFilter.cpp:159-168 (SOS of the data with itself, then sqrt) and
// Main.cpp:177 (|.|): one fused kernel; the result replaces the data as a
// single complex channel (value, 0).
template<typename T>
void ImageFilter<T>::SOS(ImageSize reconSize, const float *deapodDevice)
{
    auto &d = m_associatedData;
    if (d.dim() == 2 && reconSize.z != 1) {
        std::cerr << "Error: ImageData wrong size!" << std::endl;
        exit(1);
    }
    ImageData<T> img(d.dim(), reconSize);
    const std::size_t n = img.dataSize();
    fr_grids *g = d.deviceGrids();
    float *out = nullptr;
    check(fr_malloc(ctx(), sizeof(float) * n, (void **)&out), "allocate image");
    if (d.fftPending() != nullptr) {
        // transform + shift + crop + SOS in one call; the grids are consumed
        check(fr_fft_sos(d.fftPending(), g, reconSize.x, reconSize.y, reconSize.z, d.shiftPending() ? 1 : 0,
                         deapodDevice, out), "FFT + SOS");
        d.setFftPending(nullptr);
    } else {
        check(fr_sos(ctx(), g, reconSize.x, reconSize.y, reconSize.z, d.shiftPending() ? 1 : 0, deapodDevice, out),
              "SOS");
    }
    std::vector<float> mag(n);
    check(fr_memcpy_d2h(ctx(), mag.data(), out, sizeof(float) * n), "download image");
    fr_free(ctx(), out);
    ComplexVector<T> v(n);
    for (std::size_t i = 0; i < n; ++i) v[i] = std::complex<T>(mag[i], 0);
    img.addChannelImage(std::move(v));
    m_associatedData = std::move(img);
}

// ImageFilter.cpp:170-209: the PILS combine sum_c img_c * conj(map_c) on the
// crop (no square root).
template<typename T>
void ImageFilter<T>::SOS(const ImageData<T> &map, ImageSize reconSize)
{
    auto &d = m_associatedData;
    auto &m = const_cast<ImageData<T> &>(map);
    if (d.dim() == 2 && reconSize.z != 1) {
        std::cerr << "Error: ImageData wrong size!" << std::endl;
        exit(1);
    }
    const fr_grids *g = d.deviceGrids();
    const fr_grids *mg = m.deviceGrids();
    d.materializeFft();
    m.materializeFft();
    if (d.shiftPending() != m.shiftPending()) {
        d.materializeShift();
        m.materializeShift();
    }
    ImageData<T> img(d.dim(), reconSize);
    const std::size_t n = img.dataSize();
    void *out = nullptr;
    check(fr_malloc(ctx(), sizeof(std::complex<T>) * n, &out), "allocate image");
    check(fr_sos_map(ctx(), g, mg, reconSize.x, reconSize.y, reconSize.z, d.shiftPending() ? 1 : 0, out),
          "SOS(map)");
    ComplexVector<T> v(n);
    check(fr_memcpy_d2h(ctx(), v.data(), out, sizeof(std::complex<T>) * n), "download image");
    fr_free(ctx(), out);
    img.addChannelImage(std::move(v));
    m_associatedData = std::move(img);
}

// ImageFilter.cpp:211-220 / cuImageFilter.cpp:23-38: one batched plan
// over all channels; FFTW_BACKWARD (+1) is the reference default.
template<typename T>
void ImageFilter<T>::fftPlan(int sign)
{
    m_fft.reset();
    auto &d = m_associatedData;
    if (d.imageSize().x != d.imageSize().y || (d.dim() == 3 && d.imageSize().z != d.imageSize().x)) {
        std::cerr << "Error: FFT wrong image size" << std::endl;
        exit(1);
    }
    m_sign = sign >= 0 ? +1 : -1;
    fr_fft *f = nullptr;
    check(fr_fft_create(ctx(), d.dim(), d.imageSize().x, d.channels(), m_sign, &f), "FFT plan");
    m_fft = std::shared_ptr<fr_fft>(f, fr_fft_destroy);
}

// ImageFilter.cpp:222-229, FFT.cpp:54-91, cuFFT.cpp:24-37: in place,
// unnormalised, all channels in one execution.  Recorded as pending: the
// next reader runs it (SOS fuses it with the crop).
template<typename T>
void ImageFilter<T>::fftExecute()
{
    if (m_fft == nullptr)
        fftPlan(m_sign);
    m_associatedData.deviceGrids();
    m_associatedData.materializeShift();
    m_associatedData.setFftPending(m_fft);
}

// ImageFilter.cpp:231-251 (kept for subclasses; fftShift() itself runs on
// the device).
template<typename T>
void ImageFilter<T>::fftShift2(ComplexVector<T> *data)
{
    auto size = m_associatedData.imageSize();
    auto n0h = size.x / 2;
    auto n1h = size.y / 2;
    for (auto y = 0ul; y < n1h; y++) {
        auto y1 = y + n1h;
        for (auto x = 0ul; x < size.x; x++) {
            auto x1 = x < n0h ? x + n0h : x - n0h;
            std::swap((*data)[y * size.x + x], (*data)[y1 * size.x + x1]);
        }
    }
}

// ImageFilter.cpp:253-280
template<typename T>
void ImageFilter<T>::fftShift3(ComplexVector<T> *data)
{
    auto size = m_associatedData.imageSize();
    auto n0h = size.x / 2, n1h = size.y / 2, n2h = size.z / 2;
    for (auto z = 0ul; z < n2h; z++) {
        auto z1 = z + n2h;
        for (auto y = 0ul; y < size.y; y++) {
            auto y1 = y < n1h ? y + n1h : y - n1h;
            for (auto x = 0ul; x < size.x; x++) {
                auto x1 = x < n0h ? x + n0h : x - n0h;
                std::swap((*data)[z * size.x * size.y + y * size.x + x],
                          (*data)[z1 * size.x * size.y + y1 * size.x + x1]);
            }
        }
    }
}

template class ImageFilter<float>;
