// Non-Cartesian samples: trajectory, DCF and coil data
// (reference ReconData/ReconData.h:12-65).
//
// Same interface as the reference; the Qt-typed loaders are kept behind
// FASTRECON_WITH_QT and std:: overloads with the same semantics are added.
// On this backend the data is mirrored into HBM (one fr_samples handle) when
// a GridLut plans it; updateSingleAcquisition keeps the mirror current.
#ifndef RECONDATA_H
#define RECONDATA_H

#include <complex>
#include <memory>
#include <string>
#include <vector>
#ifdef FASTRECON_WITH_QT
#include <QStringList>
#include <QVector>
#endif

#include "common.h"

struct fr_samples;

template<typename T>
struct Point {
    T x[4];
};

template<typename T>
class ReconData
{
protected:
    ReconData(int samples, int acquisitions);

public:
    typedef std::vector<Point<T>> TrajVector;
    static std::shared_ptr<ReconData<T>> Create(int samples, int acquisitions, bool gpu = false);
    virtual ~ReconData();

    void loadFromFiles(const std::vector<std::string> &dataFileList, const std::vector<std::string> &trajFileList,
                       const std::string &dcfFileName);
    void addChannelData(ComplexVector<T> &&data);
    void setDcf(std::vector<T> &dcf);
    void loadTraj(const std::vector<T> &traj, int dim);
#ifdef FASTRECON_WITH_QT
    void loadFromFiles(const QStringList &dataFileList, const QStringList &trajFileList, const QString &dcfFileName);
    void loadTraj(const QVector<T> &traj, int dim);
#endif
    virtual void updateSingleAcquisition(const std::complex<T> *data, int acquisition, int channel = 0);
    void normalizeTraj(unsigned size);
    std::pair<T, T> getCompBounds(int comp) const {
        return m_bounds[comp];
    }
    int dataSize() const {
        return m_size;
    }
    int rcDim() const { return m_dim; }
    const TrajVector &getTraj() const {
        return m_traj;
    }
    const std::vector<T> &getDcf() const {
        return m_dcf;
    }
    const ComplexVector<T> *getChannelData(int channel) const;
    int channels() const { return m_kDataMultiChannel.size(); }
    virtual void clear();

    // ---- MI355X backend (not in the reference interface) ----
    // The HBM mirror of trajectory (as normalised now), DCF and all coils,
    // created or refreshed on demand.
    fr_samples *deviceSamples();
    // Forget the mirror (the next deviceSamples() uploads again).
    void invalidateDevice();
    // Replace the DCF from HBM (e.g. after an on-device DCF estimation).
    void refreshDcfFromDevice();
    // Incremented by every change of trajectory, DCF or coil data (GridLut
    // keys its cached gridding pass on it).
    unsigned long long version() const { return m_version; }

protected:
    int m_samples;
    int m_acquisitions;
    std::size_t m_size;
    int m_dim = 0;
    std::vector<std::pair<T, T>> m_bounds;

    TrajVector m_traj;
    std::vector<ComplexVector<T>> m_kDataMultiChannel;
    std::vector<T> m_dcf;

private:
    virtual void transformLocalTraj(T translation, T scale);
    void storeTrajComponent(const std::vector<T> &traj_c);

    std::shared_ptr<fr_samples> m_dev;
    bool m_dev_traj = false, m_dev_dcf = false;
    unsigned long long m_version = 0;
    std::vector<char> m_dev_coil;
};
#endif // RECONDATA_H
