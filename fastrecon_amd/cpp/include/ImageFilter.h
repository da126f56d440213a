// Image-stage operations (reference ImageFilter/ImageFilter.h:9-38).
//
// Same interface as the reference.  Every operation runs on the device
// grids of the ImageData: fftExecute is one batched transform over all
// channels, deferred until the next reader; fftShift is folded into the next
// reader; SOS after fftExecute runs cropped transforms that keep only the
// reconstructed region and sum the squares over channels in their last pass.
#ifndef IMAGEFILTER_H
#define IMAGEFILTER_H

#include "ImageData.h"

#ifndef FFTW_FORWARD
#define FFTW_FORWARD (-1)
#endif
#ifndef FFTW_BACKWARD
#define FFTW_BACKWARD (+1)
#endif

struct fr_fft;

template<typename T>
class ImageFilter
{
public:
    static std::shared_ptr<ImageFilter<T>> Create(ImageData<T> &imageData);
    virtual ~ImageFilter();

    void setNumOfThreads(unsigned threads) {
        m_num_threads = threads;
    }
    void fftShift();
    virtual void lowFilter(int res);
    virtual void normalize();
    virtual void crop(const ImageSize &imageSize);
    virtual void SOS(ImageSize reconSize);
    virtual void SOS(const ImageData<T> &map, ImageSize reconSize);
    virtual void fftPlan(int sign = FFTW_BACKWARD);
    void fftExecute();

    // ---- MI355X backend (not in the reference interface) ----
    // SOS divided by a deapodisation map (rc^d floats in device memory).
    void SOS(ImageSize reconSize, const float *deapodDevice);

protected:
    ImageFilter(ImageData<T> &imageData);
    virtual void fftShift2(ComplexVector<T> *data);
    virtual void fftShift3(ComplexVector<T> *data);

    std::shared_ptr<fr_fft> m_fft;

private:
    ImageData<T> &m_associatedData;
    unsigned m_num_threads = 1;
    int m_sign = FFTW_BACKWARD;
};

#endif // IMAGEFILTER_H
