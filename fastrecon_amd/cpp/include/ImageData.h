// Multi-channel grid/image container (reference ImageData/ImageData.h:7-50).
//
// Same interface as the reference.  On this backend the channels may live in
// HBM (as a fr_grids handle) while the host vectors are stale: the accessors
// copy device data back on demand (getChannelImage), so callers that only use
// the reference interface see the same values, and device-resident
// pipelines (GridLut::execute -> ImageFilter) never leave HBM until the
// final image.
#ifndef IMAGEDATA_H
#define IMAGEDATA_H

#include <memory>
#include "common.h"

struct fr_grids;
struct fr_fft;

template<typename T>
class ImageData
{
public:
    ImageData() {}
    ImageData(const int dim, const ImageSize &imageSize);

    ImageData(const ImageData<T> &imageData);
    ImageData(ImageData<T> &&imageData);
    virtual ~ImageData();

    // Copy
    ImageData<T> &operator=(const ImageData<T> &imageData);
    // Move
    ImageData<T> &operator=(ImageData<T> &&imageData);

    int channels() const {
        return m_channels;
    }

    ImageSize imageSize() const {
        return m_size;
    }

    int dim() const {
        return m_dim;
    }

    std::size_t dataSize() const;

    void addChannelImage(ComplexVector<T> &&image);
    const ComplexVector<T> *getChannelImage(int channel = 0) const;
    ComplexVector<T> *getChannelImage(int channel = 0);

    // ---- MI355X backend (not in the reference interface) ----
    // Device copy of all channels (uploaded from the host if needed).  The
    // host vectors become stale: the caller is about to change the device
    // data.
    fr_grids *deviceGrids();
    const fr_grids *deviceGrids() const;
    // Take ownership of device grids holding `channels` channels.
    void adoptDeviceGrids(fr_grids *grids, int channels);
    bool onDevice() const { return m_dev != nullptr && m_dev_valid; }
    // A pending fftshift: the device data is unshifted and readers fold the
    // shift in (ImageFilter::SOS) or apply it before the data leaves HBM.
    bool shiftPending() const { return m_shift_pending; }
    void setShiftPending(bool pending) { m_shift_pending = pending; }
    void materializeShift();
    // A pending transform (ImageFilter::fftExecute): the device data is not
    // yet transformed; SOS fuses it with the shift and crop, every other
    // reader runs it first (before any pending shift).
    fr_fft *fftPending() const { return m_fft_pending.get(); }
    void setFftPending(std::shared_ptr<fr_fft> fft) { m_fft_pending = std::move(fft); }
    void materializeFft();

protected:
    virtual void copy(const ImageData<T> &imageData);
    virtual void move(ImageData<T> &imageData);

    int m_dim = 0;
    ImageSize m_size = {0, 0, 0};
    int m_channels = 0;

    mutable std::vector<ComplexVector<T>> m_data_multichannel;

private:
    void syncToHost() const;
    void dropDevice();

    std::shared_ptr<fr_grids> m_dev;
    mutable bool m_host_valid = true;
    bool m_dev_valid = false;
    bool m_shift_pending = false;
    std::shared_ptr<fr_fft> m_fft_pending;
};

#endif // IMAGEDATA_H
