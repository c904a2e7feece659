/* C++ mirror of the reference's H5Part file I/O on one rank (main/src/io/ifile_io.hpp,
 * ifile_io_hdf5.cpp): H5PartWriter and H5PartReader with the IFileWriter / IFileReader method names,
 * over the C ABI of sphexa_amd_io.h, and the ParticlesData attribute set of loadOrStoreAttributes
 * (sph/include/sph/particles_data.hpp:147-205) for a Dataset. Field pointers may be device or host
 * memory. Errors throw std::runtime_error, a missing attribute std::out_of_range, as the reference. */
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "sphexa_amd_io.h"

namespace sphexa_amd
{

namespace io_detail
{

//! element type of T as H5Part stores it (h5part_wrapper.hpp:56-106)
template<class T>
constexpr int h5Type()
{
    using U = std::remove_cv_t<T>;
    if constexpr (std::is_same_v<U, double>) { return SPHX_H5_F64; }
    else if constexpr (std::is_same_v<U, float>) { return SPHX_H5_F32; }
    else if constexpr (std::is_integral_v<U> && sizeof(U) == 4) { return SPHX_H5_I32; }
    else if constexpr (std::is_integral_v<U> && sizeof(U) == 8) { return SPHX_H5_I64; }
    else
    {
        static_assert(std::is_integral_v<U> && sizeof(U) == 1, "element types: double, float, 8/32/64-bit integers");
        return SPHX_H5_I8;
    }
}

inline void check(int code)
{
    if (code == SPHX_H5_OK) { return; }
    if (code == SPHX_H5_ERR_MISSING) { throw std::out_of_range(sphx_h5_last_error()); }
    throw std::runtime_error(std::string("sphexa_amd io: ") + sphx_h5_last_error());
}

inline std::vector<std::string> names(sphx_h5file_t* f, int step, int kind)
{
    int64_t n = 0;
    check(sphx_h5_count(f, step, kind, &n));
    std::vector<std::string> out;
    for (int64_t i = 0; i < n; ++i)
    {
        std::string buf(256, '\0');
        int         code;
        // a name that does not fit is an argument error: grow the buffer and ask again
        while ((code = sphx_h5_name(f, step, kind, i, buf.data(), buf.size())) == SPHX_H5_ERR_ARG &&
               buf.size() < (size_t(1) << 20))
        {
            buf.resize(buf.size() * 4);
        }
        check(code);
        out.emplace_back(buf.c_str());
    }
    return out;
}

} // namespace io_detail

//! H5PartWriter (ifile_io_hdf5.cpp:27-129) on one rank
class H5PartWriter
{
public:
    ~H5PartWriter() { closeStep(); }

    int         rank() const { return 0; }
    int         numRanks() const { return 1; }
    std::string suffix() const { return ".h5"; }

    //! start the next step of `path` with particles [firstIndex, lastIndex) of the fields written
    //! next; a new path is created, an existing one appended to
    void addStep(size_t firstIndex, size_t lastIndex, const std::string& path)
    {
        firstIndex_ = firstIndex;
        if (!f_ || path != path_)
        {
            closeStep();
            io_detail::check(sphx_h5_open(path.c_str(), SPHX_H5_APPEND, &f_));
            path_ = path;
        }
        if (lastIndex > firstIndex) { io_detail::check(sphx_h5_add_step(f_, uint64_t(lastIndex - firstIndex))); }
    }

    template<class T>
    void stepAttribute(const std::string& key, const T* val, int64_t size)
    {
        io_detail::check(sphx_h5_write_attrib(f_, 1, key.c_str(), io_detail::h5Type<T>(), val, uint64_t(size)));
    }

    template<class T>
    void fileAttribute(const std::string& key, const T* val, int64_t size)
    {
        io_detail::check(sphx_h5_write_attrib(f_, 0, key.c_str(), io_detail::h5Type<T>(), val, uint64_t(size)));
    }

    //! the step's particles of `field` (device or host), from firstIndex on
    template<class T>
    void writeField(const std::string& key, const T* field, int = 0)
    {
        io_detail::check(sphx_h5_write_field(f_, key.c_str(), io_detail::h5Type<T>(), field + firstIndex_));
    }

    void closeStep()
    {
        if (f_)
        {
            sphx_h5_close(f_);
            f_ = nullptr;
        }
    }

private:
    sphx_h5file_t* f_{nullptr};
    std::string    path_;
    size_t         firstIndex_{0};
};

//! H5PartReader (ifile_io_hdf5.cpp:131-310) on one rank (every particle of the step is local)
class H5PartReader
{
public:
    ~H5PartReader() { closeStep(); }

    int rank() const { return 0; }

    //! open `path` at `step` (negative: the last step)
    void setStep(const std::string& path, int step)
    {
        closeStep();
        io_detail::check(sphx_h5_open(path.c_str(), SPHX_H5_READ, &f_));
        int64_t steps = 0;
        io_detail::check(sphx_h5_num_steps(f_, &steps));
        if (steps == 0) { return; }
        io_detail::check(sphx_h5_set_step(f_, step, &n_));
    }

    int64_t                  numParticles() const { return int64_t(n_); }
    uint64_t                 localNumParticles() const { return n_; }
    uint64_t                 globalNumParticles() const { return n_; }
    std::vector<std::string> fileAttributes() { return io_detail::names(f_, 0, 1); }
    std::vector<std::string> stepAttributes() { return io_detail::names(f_, 1, 1); }

    int64_t fileAttributeSize(const std::string& key) { return attributeSize(0, key); }
    int64_t stepAttributeSize(const std::string& key) { return attributeSize(1, key); }

    template<class T>
    void fileAttribute(const std::string& key, T* val, int64_t size)
    {
        io_detail::check(sphx_h5_read_attrib(f_, 0, key.c_str(), io_detail::h5Type<T>(), val, uint64_t(size)));
    }

    template<class T>
    void stepAttribute(const std::string& key, T* val, int64_t size)
    {
        io_detail::check(sphx_h5_read_attrib(f_, 1, key.c_str(), io_detail::h5Type<T>(), val, uint64_t(size)));
    }

    template<class T>
    void readField(const std::string& key, T* field)
    {
        io_detail::check(sphx_h5_read_field(f_, key.c_str(), io_detail::h5Type<T>(), field));
    }

    void closeStep()
    {
        if (f_)
        {
            sphx_h5_close(f_);
            f_ = nullptr;
        }
        n_ = 0;
    }

private:
    int64_t attributeSize(int step, const std::string& key)
    {
        int      t = 0;
        uint64_t n = 0;
        io_detail::check(sphx_h5_attrib_info(f_, step, key.c_str(), &t, &n));
        return int64_t(n);
    }

    sphx_h5file_t* f_{nullptr};
    uint64_t       n_{0};
};

/*! ParticlesData::loadOrStoreAttributes (sph/include/sph/particles_data.hpp:147-205) for a Dataset
 *  with the ParticlesData scalars (DeviceDataset): with an H5PartWriter every attribute is written;
 *  with an H5PartReader the required ones (iteration, numParticlesGlobal, time, minDt, minDt_m1,
 *  gravConstant) are read and the optional ones keep their value when the file lacks them. */
template<class Archive, class Dataset>
void loadOrStoreAttributes(Archive* ar, Dataset& d)
{
    constexpr bool loading = std::is_same_v<Archive, H5PartReader>;
    auto io = [ar](const std::string& name, auto* location, bool optional)
    {
        try
        {
            ar->stepAttribute(name, location, 1);
        }
        catch (const std::out_of_range&)
        {
            if (!optional || !loading) { throw; }
        }
    };
    io("iteration", &d.iteration, false);
    io("numParticlesGlobal", &d.numParticlesGlobal, false);
    io("ng0", &d.ng0, true);
    io("ngmax", &d.ngmax, true);
    io("time", &d.ttot, false);
    io("minDt", &d.minDt, false);
    io("minDt_m1", &d.minDt_m1, false);
    io("Kcour", &d.Kcour, true);
    io("Krho", &d.Krho, true);
    io("gravConstant", &d.g, false);
    io("eps", &d.eps, true);
    io("etaAcc", &d.etaAcc, true);
    io("gamma", &d.gamma, true);
    io("eosChoice", &d.eosChoice, true);
    io("muiConst", &d.muiConst, true);
    io("soundSpeedConst", &d.soundSpeedConst, true);
    io("polytropic_index", &d.polytropic_index, true);
    io("polytropic_const", &d.polytropic_const, true);
    io("alphamin", &d.alphamin, true);
    io("alphamax", &d.alphamax, true);
    io("decay_constant", &d.decay_constant, true);
    io("sincIndex", &d.sincIndex, true);
}

} // namespace sphexa_amd
