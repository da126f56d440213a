// ReconData (reference ReconData/ReconData.cpp:12-237) with an HBM mirror.
// Host-side behaviour, messages and exit codes follow the reference line by
// line; the mirror (fr_samples) is refreshed lazily by deviceSamples().
#include <algorithm>
#include <fstream>
#include <iostream>
#include <limits>

#include "ReconData.h"
#include "hip_backend.h"

using fastrecon::check;
using fastrecon::ctx;

template<typename T>
ReconData<T>::ReconData(int samples, int acquisitions) : m_samples(samples), m_acquisitions(acquisitions)
{
    m_size = (std::size_t)samples * (std::size_t)acquisitions;
}

template<typename T>
ReconData<T>::~ReconData() {}

// ReconData.cpp:16-31.  This library's only backend is the MI355X one, so
// both values of `gpu` give the device-backed object.
template<typename T>
std::shared_ptr<ReconData<T>> ReconData<T>::Create(int samples, int acquisitions, bool gpu)
{
    (void)gpu;
    return std::shared_ptr<ReconData<T>>(new ReconData<T>(samples, acquisitions));
}

template<typename T>
void ReconData<T>::addChannelData(ComplexVector<T> &&data)
{
    if (m_size != data.size())                          // ReconData.cpp:35-40
    {
        std::cerr << "Error: trajectory and data have different size!" << std::endl;
        exit(1);
    }
    m_kDataMultiChannel.push_back(std::move(data));
    m_dev.reset();                                      // coil count changed
    ++m_version;
}

template<typename T>
void ReconData<T>::setDcf(std::vector<T> &dcf)
{
    if (m_size != dcf.size())                           // ReconData.cpp:47-52
    {
        std::cerr << "Error: data size does not match!" << std::endl;
        exit(1);
    }
    m_dcf = std::move(dcf);
    m_dev_dcf = false;
    ++m_version;
}

template<typename T>
void ReconData<T>::storeTrajComponent(const std::vector<T> &traj_c)
{
    if (m_size != traj_c.size())                        // ReconData.cpp:58-63
    {
        std::cerr << "Error: data size does not match!" << std::endl;
        exit(1);
    }
    if (m_dim == 4)
    {
        std::cout << "4-D trajectories have been loaded, ignoring additional data." << std::endl;
        return;
    }
    m_traj.resize(m_size);

    auto bound = std::minmax_element(traj_c.begin(), traj_c.end());
    m_bounds.push_back(std::make_pair(*bound.first, *bound.second));
    std::cout << "Range: " << '(' << *bound.first << ", " << *bound.second << ')' << std::endl;

    auto in = traj_c.begin();
    for (auto &p : m_traj)
        p.x[m_dim] = *(in++);
    ++m_dim;
    m_dev_traj = false;
    ++m_version;
}

template<typename T>
static bool read_exact(const std::string &name, void *dst, std::size_t bytes)
{
    std::ifstream f(name, std::ios::binary);
    if (!f) return false;
    f.read(static_cast<char *>(dst), (std::streamsize)bytes);
    return (std::size_t)f.gcount() == bytes;
}

template<typename T>
void ReconData<T>::loadFromFiles(const std::vector<std::string> &dataFileList,
                                 const std::vector<std::string> &trajFileList, const std::string &dcfFileName)
{
    std::cout << std::endl << "Read trajectory:" << std::endl;       // ReconData.cpp:83-144
    m_dim = 0;
    m_bounds.clear();

    for (const auto &name : trajFileList)
    {
        std::cout << name << std::endl;
        std::vector<T> traj_c(m_size);
        if (!read_exact<T>(name, traj_c.data(), m_size * sizeof(T)))
        {
            std::cout << "Error: wrong data size in " << name << std::endl;
            std::exit(1);
        }
        storeTrajComponent(traj_c);
    }

    std::cout << std::endl << "Read dcf:" << std::endl;
    std::cout << dcfFileName << std::endl;
    m_dcf.resize(m_size);
    if (!read_exact<T>(dcfFileName, m_dcf.data(), m_size * sizeof(T)))
    {
        std::cout << "Error: wrong data size in " << dcfFileName << std::endl;
        std::exit(1);
    }
    m_dev_dcf = false;
    ++m_version;

    std::cout << std::endl << "Read data:" << std::endl;
    ComplexVector<T> kdata;
    for (const auto &name : dataFileList)
    {
        kdata.resize(m_size);
        std::cout << name << std::endl;
        if (!read_exact<T>(name, kdata.data(), m_size * sizeof(typename ComplexVector<T>::value_type)))
        {
            std::cout << "Error: wrong data size in " << name << std::endl;
            std::exit(1);
        }
        addChannelData(std::move(kdata));
    }
}

template<typename T>
void ReconData<T>::loadTraj(const std::vector<T> &traj, int dim)
{
    if ((std::size_t)traj.size() != m_size * (dim + 1)) {        // ReconData.cpp:146-177
        std::cerr << "Error: data size does not match!" << std::endl << std::flush;
        return;
    }
    m_dim = dim;
    m_traj.resize(m_size);
    m_dcf.resize(m_size);
    m_bounds.resize(dim);

    for (auto &bound : m_bounds) {
        bound.first = std::numeric_limits<T>::max();
        bound.second = std::numeric_limits<T>::lowest();
    }

    auto in = traj.begin();
    auto it_dcf = m_dcf.begin();
    for (auto &p : m_traj)
    {
        for (auto i = 0; i < dim; i++) {
            p.x[i] = *(in++);
            if (m_bounds[i].first > p.x[i])
                m_bounds[i].first = p.x[i];
            if (m_bounds[i].second < p.x[i])
                m_bounds[i].second = p.x[i];
        }
        *(it_dcf++) = *(in++);
    }
    m_dev_traj = false;
    m_dev_dcf = false;
    ++m_version;
}

#ifdef FASTRECON_WITH_QT
template<typename T>
void ReconData<T>::loadFromFiles(const QStringList &dataFileList, const QStringList &trajFileList,
                                 const QString &dcfFileName)
{
    std::vector<std::string> d, t;
    for (const auto &s : dataFileList) d.push_back(s.toStdString());
    for (const auto &s : trajFileList) t.push_back(s.toStdString());
    loadFromFiles(d, t, dcfFileName.toStdString());
}

template<typename T>
void ReconData<T>::loadTraj(const QVector<T> &traj, int dim)
{
    loadTraj(std::vector<T>(traj.begin(), traj.end()), dim);
}
#endif

template<typename T>
void ReconData<T>::updateSingleAcquisition(const std::complex<T> *data, int acquisition, int channel)
{
    if (channel > (int)m_kDataMultiChannel.size() - 1)           // ReconData.cpp:179-189
    {
        std::cerr << "Error: channel " << channel << " is empty!" << std::endl << std::flush;
        return;
    }
    auto &channel_data = m_kDataMultiChannel[channel];
    std::copy_n(data, m_samples, channel_data.begin() + (std::size_t)m_samples * acquisition);
    ++m_version;
    // cuReconData::updateSingleAcquisition (cuReconData.cu:52-62): keep the mirror current
    if (m_dev && channel < (int)m_dev_coil.size() && m_dev_coil[channel])
        check(fr_samples_update(m_dev.get(), channel, (int64_t)m_samples * acquisition, m_samples, data),
              "update acquisition");
}

template<typename T>
const ComplexVector<T> *ReconData<T>::getChannelData(int channel) const
{
    if (channel < 0 || channel > (int)m_kDataMultiChannel.size() - 1)
        return nullptr;
    return &m_kDataMultiChannel[channel];
}

template<typename T>
void ReconData<T>::transformLocalTraj(T translation, T scale)
{
    for (auto &sample : m_traj)                                   // ReconData.cpp:201-209
        for (int comp = 0; comp < m_dim; ++comp)
            sample.x[comp] = (sample.x[comp] + translation) * scale;
    m_dev_traj = false;
    ++m_version;
}

template<typename T>
void ReconData<T>::normalizeTraj(unsigned size)
{
    auto bounds = getCompBounds(0);                               // ReconData.cpp:211-225
    auto tr = -bounds.first;
    auto scale = (size - 1) / (bounds.second - bounds.first);

    for (int comp = 0; comp < m_dim; ++comp)
    {
        m_bounds[comp].first = (m_bounds[comp].first + tr) * scale;
        m_bounds[comp].second = (m_bounds[comp].second + tr) * scale;
    }
    transformLocalTraj(tr, scale);
}

template<typename T>
void ReconData<T>::clear()
{
    m_size = 0;
    m_traj.clear();
    m_dcf.clear();
    m_kDataMultiChannel.clear();
    m_bounds.clear();
    m_dev.reset();
    ++m_version;
}

template<typename T>
fr_samples *ReconData<T>::deviceSamples()
{
    const int C = channels();
    if (m_dim != 2 && m_dim != 3) {
        std::cerr << "Error: gridding needs 2 or 3 trajectory components (have " << m_dim << ")" << std::endl;
        exit(1);
    }
    if (C < 1) {
        std::cerr << "Error: no channel data" << std::endl;
        exit(1);
    }
    if (!m_dev) {
        fr_samples *s = nullptr;
        check(fr_samples_create(ctx(), m_dim, (int64_t)m_size, C, FR_SAMPLES_C64, &s), "allocate device samples");
        m_dev = std::shared_ptr<fr_samples>(s, fr_samples_destroy);
        m_dev_traj = m_dev_dcf = false;
        m_dev_coil.assign(C, 0);
    }
    if (!m_dev_traj) {
        check(fr_samples_set_traj(m_dev.get(), &m_traj[0].x[0], 4), "upload trajectory");
        m_dev_traj = true;
    }
    if (!m_dev_dcf) {
        if (m_dcf.size() != m_size) {
            std::cerr << "Error: data size does not match!" << std::endl;
            exit(1);
        }
        check(fr_samples_set_dcf(m_dev.get(), m_dcf.data()), "upload dcf");
        m_dev_dcf = true;
    }
    for (int c = 0; c < C; ++c)
        if (!m_dev_coil[c]) {
            check(fr_samples_set_coil(m_dev.get(), c, m_kDataMultiChannel[c].data()), "upload channel");
            m_dev_coil[c] = 1;
        }
    return m_dev.get();
}

template<typename T>
void ReconData<T>::invalidateDevice()
{
    m_dev.reset();
}

template<typename T>
void ReconData<T>::refreshDcfFromDevice()
{
    if (!m_dev) return;
    m_dcf.resize(m_size);
    check(fr_samples_get_dcf(m_dev.get(), m_dcf.data()), "download dcf");
    m_dev_dcf = true;
    ++m_version;
}

template class ReconData<float>;
