// Particle fields and attributes in the H5Part layout of HDF5 files (include/sphexa_amd_io.h), written
// with the HDF5 C API directly. The layout is the one extern/h5part/H5Part.c produces and reads:
// groups "Step#<k>" under the root (H5Part.c:600-626, step names without zero padding), 1-D datasets
// of native types per field (H5Part.c:920-1000), 1-D attributes on the step group or on the root
// group (H5Part.c:1330-1375). The reference drives it through H5PartWriter / H5PartReader
// (main/src/io/ifile_io_hdf5.cpp:27-310) and the typed helpers of h5part_wrapper.hpp.
#include <hip/hip_runtime_api.h>
#include <hdf5.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "sphexa_amd_io.h"

struct sphx_h5file
{
    hid_t    file = -1;
    hid_t    step = -1; // group of the current step
    int      mode = 0;
    uint64_t n    = 0;  // particles of the current step
    void*       pinned      = nullptr; // page-locked staging of device fields: two chunk buffers
    size_t      pinnedBytes = 0;
    hipStream_t copies      = nullptr; // stream of the staging copies (default priority)
    hipEvent_t  done[2]     = {nullptr, nullptr}; // copy of the chunk in staging buffer 0 / 1 complete
};

namespace
{

thread_local std::string g_error;

struct Failure
{
    int         code;
    std::string what;
};

[[noreturn]] void fail(int code, std::string what) { throw Failure{code, std::move(what)}; }

template<class F>
int guarded(F&& body)
{
    try
    {
        body();
        g_error.clear();
        return SPHX_H5_OK;
    }
    catch (const Failure& e)
    {
        g_error = e.what;
        return e.code;
    }
    catch (const std::exception& e)
    {
        g_error = e.what();
        return SPHX_H5_ERR_IO;
    }
}

//! native HDF5 type of an element type (H5PART_FLOAT64 ... H5PART_CHAR, H5Part.h:58-63)
hid_t nativeType(int type)
{
    switch (type)
    {
        case SPHX_H5_F64: return H5T_NATIVE_DOUBLE;
        case SPHX_H5_F32: return H5T_NATIVE_FLOAT;
        case SPHX_H5_I32: return H5T_NATIVE_INT32;
        case SPHX_H5_I64: return H5T_NATIVE_INT64;
        case SPHX_H5_I8: return H5T_NATIVE_CHAR;
        default: fail(SPHX_H5_ERR_ARG, "unknown element type " + std::to_string(type));
    }
}

size_t typeSize(int type)
{
    static const size_t sizes[] = {8, 4, 4, 8, 1};
    return sizes[type];
}

//! element type of a stored HDF5 type (the first of H5PartTypes it matches, h5part_wrapper.hpp:47)
int typeOf(hid_t t)
{
    const H5T_class_t c = H5Tget_class(t);
    const size_t      s = H5Tget_size(t);
    if (c == H5T_FLOAT && s == 8) { return SPHX_H5_F64; }
    if (c == H5T_FLOAT && s == 4) { return SPHX_H5_F32; }
    if (c == H5T_INTEGER && s == 4) { return SPHX_H5_I32; }
    if (c == H5T_INTEGER && s == 8) { return SPHX_H5_I64; }
    if (c == H5T_INTEGER && s == 1) { return SPHX_H5_I8; }
    fail(SPHX_H5_ERR_TYPE, "unsupported element type in file");
}

void check(herr_t r, const std::string& what)
{
    if (r < 0) { fail(SPHX_H5_ERR_IO, "HDF5: " + what); }
}

//! name of link or attribute i of a group, whatever its length (`get` is H5Lget_name_by_idx or
//! H5Aget_name_by_idx: called without a buffer it returns the length)
template<class Get>
std::string nameAt(Get get, hid_t group, hsize_t i)
{
    ssize_t len = get(group, ".", H5_INDEX_NAME, H5_ITER_INC, i, nullptr, 0, H5P_DEFAULT);
    if (len < 0) { fail(SPHX_H5_ERR_IO, "HDF5: name of entry " + std::to_string(i)); }
    std::string name(size_t(len) + 1, '\0');
    check(get(group, ".", H5_INDEX_NAME, H5_ITER_INC, i, name.data(), name.size(), H5P_DEFAULT), "name");
    name.resize(size_t(len));
    return name;
}

//! closes an HDF5 object at scope exit
struct Handle
{
    hid_t id;
    herr_t (*close)(hid_t);
    ~Handle()
    {
        if (id >= 0) { close(id); }
    }
};

bool onDevice(const void* p)
{
    hipPointerAttribute_t a{};
    bool                  dev = hipPointerGetAttributes(&a, p) == hipSuccess && a.type == hipMemoryTypeDevice;
    (void)hipGetLastError(); // host memory unknown to HIP leaves an error to clear
    return dev;
}

//! device fields move in chunks of this many bytes, through two page-locked staging buffers, so
//! the copy of one chunk overlaps HDF5's work on the other
constexpr size_t kChunkBytes = size_t(16) << 20;

//! two page-locked chunk buffers (each `chunk` bytes), the copy stream and its events; false when
//! page-locked memory or the stream cannot be had (the caller then stages through pageable memory)
bool pinnedStaging(sphx_h5file* f, size_t chunk)
{
    if (f->copies == nullptr)
    {
        if (hipStreamCreate(&f->copies) != hipSuccess)
        {
            f->copies = nullptr;
            (void)hipGetLastError();
            return false;
        }
        for (auto& e : f->done)
        {
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            {
                e = nullptr;
                (void)hipGetLastError();
                return false;
            }
        }
    }
    if (f->done[0] == nullptr || f->done[1] == nullptr) { return false; }
    if (2 * chunk > f->pinnedBytes)
    {
        if (f->pinned) { (void)hipHostFree(f->pinned); }
        f->pinned      = nullptr;
        f->pinnedBytes = 0;
        if (hipHostMalloc(&f->pinned, 2 * chunk, hipHostMallocDefault) == hipSuccess) { f->pinnedBytes = 2 * chunk; }
        else
        {
            f->pinned = nullptr;
            (void)hipGetLastError();
        }
    }
    return f->pinned != nullptr;
}

void releaseStaging(sphx_h5file* f)
{
    if (f->copies) { (void)hipStreamSynchronize(f->copies); }
    for (auto& e : f->done)
    {
        if (e) { (void)hipEventDestroy(e); }
        e = nullptr;
    }
    if (f->copies) { (void)hipStreamDestroy(f->copies); }
    f->copies = nullptr;
    if (f->pinned) { (void)hipHostFree(f->pinned); }
    f->pinned      = nullptr;
    f->pinnedBytes = 0;
}

//! elements [off, off + cnt) of dataset d from or into host memory `buf` (H5Dwrite / H5Dread on a
//! hyperslab of the 1-D file space)
void slab(hid_t d, hid_t t, hsize_t off, hsize_t cnt, void* buf, bool write, const std::string& name)
{
    Handle fs{H5Dget_space(d), H5Sclose};
    check(H5Sselect_hyperslab(fs.id, H5S_SELECT_SET, &off, nullptr, &cnt, nullptr), "select " + name);
    Handle ms{H5Screate_simple(1, &cnt, nullptr), H5Sclose};
    if (write) { check(H5Dwrite(d, t, ms.id, fs.id, H5P_DEFAULT, buf), "write " + name); }
    else { check(H5Dread(d, t, ms.id, fs.id, H5P_DEFAULT, buf), "read " + name); }
}

//! waits for the staging copies however the caller leaves (an HDF5 error mid-field included)
struct CopiesDrained
{
    hipStream_t s;
    ~CopiesDrained()
    {
        if (s) { (void)hipStreamSynchronize(s); }
    }
};

/*! n elements of size es between device memory `dev` and dataset d, in chunks through the two
 *  page-locked buffers: writing, the device-to-host copy of chunk c + 1 runs while HDF5 writes
 *  chunk c; reading, the host-to-device copy of chunk c runs while HDF5 reads chunk c + 1. */
void pipelined(sphx_h5file* f, hid_t d, hid_t t, size_t es, uint64_t n, void* dev, bool write, const std::string& name)
{
    const uint64_t ce      = std::max<uint64_t>(1, kChunkBytes / es); // elements per chunk
    const uint64_t nChunks = (n + ce - 1) / ce;
    char*          pin     = static_cast<char*>(f->pinned);
    char*          buf[2]  = {pin, pin + (nChunks > 1 ? ce * es : 0)}; // one chunk: buffer 0 only
    char*          base    = static_cast<char*>(dev);
    CopiesDrained  drain{f->copies};
    auto           count   = [&](uint64_t c) { return std::min<uint64_t>(ce, n - c * ce); };
    auto           hipOk   = [&](hipError_t e, const char* what) {
        if (e != hipSuccess)
        {
            (void)hipGetLastError();
            fail(SPHX_H5_ERR_IO, std::string(what) + " of field " + name);
        }
    };
    // the copy stream is ordered after no other stream: finish the device work that produced or
    // still uses the field first, as a synchronous copy would
    hipOk(hipDeviceSynchronize(), "device synchronisation");
    if (write)
    {
        hipOk(hipMemcpyAsync(buf[0], base, count(0) * es, hipMemcpyDeviceToHost, f->copies), "device to host copy");
        hipOk(hipEventRecord(f->done[0], f->copies), "event");
        for (uint64_t c = 0; c < nChunks; ++c)
        {
            const int b = int(c & 1);
            if (c + 1 < nChunks)
            {
                // buffer b ^ 1 held chunk c - 1, whose HDF5 write has returned
                hipOk(hipMemcpyAsync(buf[b ^ 1], base + (c + 1) * ce * es, count(c + 1) * es, hipMemcpyDeviceToHost,
                                     f->copies),
                      "device to host copy");
                hipOk(hipEventRecord(f->done[b ^ 1], f->copies), "event");
            }
            hipOk(hipEventSynchronize(f->done[b]), "device to host copy");
            slab(d, t, c * ce, count(c), buf[b], true, name);
        }
    }
    else
    {
        for (uint64_t c = 0; c < nChunks; ++c)
        {
            const int b = int(c & 1);
            // buffer b held chunk c - 2: its host-to-device copy must be done before HDF5 refills it
            if (c >= 2) { hipOk(hipEventSynchronize(f->done[b]), "host to device copy"); }
            slab(d, t, c * ce, count(c), buf[b], false, name);
            hipOk(hipMemcpyAsync(base + c * ce * es, buf[b], count(c) * es, hipMemcpyHostToDevice, f->copies),
                  "host to device copy");
            hipOk(hipEventRecord(f->done[b], f->copies), "event");
        }
        hipOk(hipStreamSynchronize(f->copies), "host to device copy");
    }
}

std::string stepName(int64_t k) { return "Step#" + std::to_string(k); }

//! the dataset name H5Part stores for a field name: names over 64 characters keep their first 63
//! (_normalize_dataset_name, H5Part.c:906-921, with H5PART_DATANAME_LEN = 64), on writing and on
//! reading alike
std::string datasetName(const char* name)
{
    std::string s(name);
    if (s.size() > 64) { s.resize(63); }
    return s;
}

//! "Step#<digits>" groups of the root (H5PartGetNumSteps)
int64_t countSteps(hid_t file)
{
    H5G_info_t info;
    check(H5Gget_info(file, &info), "group info of /");
    int64_t steps = 0;
    for (hsize_t i = 0; i < info.nlinks; ++i)
    {
        const std::string name = nameAt(H5Lget_name_by_idx, file, i);
        if (name.rfind("Step#", 0) == 0 && name.size() > 5 &&
            name.find_first_not_of("0123456789", 5) == std::string::npos)
        {
            ++steps;
        }
    }
    return steps;
}

sphx_h5file* handle(sphx_h5file_t* f)
{
    if (f == nullptr || f->file < 0) { fail(SPHX_H5_ERR_ARG, "file not open"); }
    return f;
}

hid_t stepGroup(sphx_h5file* f)
{
    if (f->step < 0) { fail(SPHX_H5_ERR_ARG, "no current step (add_step or set_step first)"); }
    return f->step;
}

//! the group an attribute call addresses: the current step (step != 0) or the root
hid_t attribOwner(sphx_h5file* f, int step) { return step ? stepGroup(f) : f->file; }

//! names of the datasets (kind 0) or attributes (kind 1) of a group, in name order
std::vector<std::string> names(hid_t group, int kind)
{
    std::vector<std::string> out;
    if (kind == 0)
    {
        H5G_info_t info;
        check(H5Gget_info(group, &info), "group info");
        for (hsize_t i = 0; i < info.nlinks; ++i)
        {
            std::string name = nameAt(H5Lget_name_by_idx, group, i);
            H5O_info_t  oi;
            if (H5Oget_info_by_name(group, name.c_str(), &oi, H5P_DEFAULT) >= 0 && oi.type == H5O_TYPE_DATASET)
                out.push_back(std::move(name));
        }
    }
    else
    {
        H5O_info_t oi;
        check(H5Oget_info(group, &oi), "object info");
        for (hsize_t i = 0; i < oi.num_attrs; ++i)
            out.push_back(nameAt(H5Aget_name_by_idx, group, i));
    }
    return out;
}

//! count values of element type S at src as element type `to` into dst
template<class S>
void convertFrom(const S* src, int to, void* dst, size_t count)
{
    for (size_t i = 0; i < count; ++i)
    {
        switch (to)
        {
            case SPHX_H5_F64: static_cast<double*>(dst)[i] = static_cast<double>(src[i]); break;
            case SPHX_H5_F32: static_cast<float*>(dst)[i] = static_cast<float>(src[i]); break;
            case SPHX_H5_I32: static_cast<int32_t*>(dst)[i] = static_cast<int32_t>(src[i]); break;
            case SPHX_H5_I64: static_cast<int64_t*>(dst)[i] = static_cast<int64_t>(src[i]); break;
            default: static_cast<signed char*>(dst)[i] = static_cast<signed char>(src[i]); break;
        }
    }
}

//! count values of element type `from` as element type `to` (the conversions of std::copy)
void convert(const void* src, int from, void* dst, int to, size_t count)
{
    switch (from)
    {
        case SPHX_H5_F64: convertFrom(static_cast<const double*>(src), to, dst, count); break;
        case SPHX_H5_F32: convertFrom(static_cast<const float*>(src), to, dst, count); break;
        case SPHX_H5_I32: convertFrom(static_cast<const int32_t*>(src), to, dst, count); break;
        case SPHX_H5_I64: convertFrom(static_cast<const int64_t*>(src), to, dst, count); break;
        default: convertFrom(static_cast<const signed char*>(src), to, dst, count); break;
    }
}

/*! readAttribute's rule (h5part_wrapper.hpp:198-220): both floating point, or the requested type is
 *  the common type of both (ranks char < int < int64 < float < double, as std::common_type orders
 *  them) */
bool convertible(int from, int to)
{
    static const int rank[] = {4, 3, 1, 2, 0}; // F64, F32, I32, I64, I8
    const bool       bothFloating = from <= SPHX_H5_F32 && to <= SPHX_H5_F32;
    return bothFloating || rank[to] >= rank[from];
}

} // namespace

extern "C" {

const char* sphx_h5_last_error(void) { return g_error.c_str(); }

int sphx_h5_open(const char* path, int mode, sphx_h5file_t** out)
{
    return guarded([&] {
        if (path == nullptr || out == nullptr) { fail(SPHX_H5_ERR_ARG, "null argument"); }
        *out = nullptr;
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr); // errors are returned, not printed
        hid_t file = -1;
        if (mode == SPHX_H5_READ) { file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT); }
        else if (mode == SPHX_H5_WRITE) { file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT); }
        else if (mode == SPHX_H5_APPEND)
        {
            file = H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
            if (file < 0) { file = H5Fcreate(path, H5F_ACC_EXCL, H5P_DEFAULT, H5P_DEFAULT); }
        }
        else { fail(SPHX_H5_ERR_ARG, "mode must be SPHX_H5_READ, SPHX_H5_WRITE or SPHX_H5_APPEND"); }
        if (file < 0) { fail(mode == SPHX_H5_READ ? SPHX_H5_ERR_MISSING : SPHX_H5_ERR_IO, std::string("cannot open ") + path); }
        auto* f = new sphx_h5file;
        f->file = file;
        f->mode = mode;
        *out    = f;
    });
}

int sphx_h5_close(sphx_h5file_t* f)
{
    return guarded([&] {
        if (f == nullptr) { return; }
        if (f->step >= 0) { H5Gclose(f->step); }
        herr_t r = f->file >= 0 ? H5Fclose(f->file) : 0;
        releaseStaging(f);
        delete f;
        check(r, "close");
    });
}

int sphx_h5_num_steps(sphx_h5file_t* f, int64_t* out)
{
    return guarded([&] {
        if (out == nullptr) { fail(SPHX_H5_ERR_ARG, "null argument"); }
        *out = countSteps(handle(f)->file);
    });
}

int sphx_h5_add_step(sphx_h5file_t* f, uint64_t n)
{
    return guarded([&] {
        handle(f);
        if (f->mode == SPHX_H5_READ) { fail(SPHX_H5_ERR_ARG, "file opened for reading"); }
        if (f->step >= 0)
        {
            H5Gclose(f->step);
            f->step = -1;
        }
        const std::string name = stepName(countSteps(f->file));
        f->step                = H5Gcreate2(f->file, name.c_str(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (f->step < 0) { fail(SPHX_H5_ERR_IO, "cannot create " + name); }
        f->n = n;
    });
}

int sphx_h5_set_step(sphx_h5file_t* f, int64_t step, uint64_t* n)
{
    return guarded([&] {
        handle(f);
        const int64_t steps = countSteps(f->file);
        if (step < 0) { step = steps - 1; }
        if (step < 0 || step >= steps) { fail(SPHX_H5_ERR_MISSING, "no step " + std::to_string(step)); }
        if (f->step >= 0)
        {
            H5Gclose(f->step);
            f->step = -1;
        }
        const std::string name = stepName(step);
        f->step                = H5Gopen2(f->file, name.c_str(), H5P_DEFAULT);
        if (f->step < 0) { fail(SPHX_H5_ERR_MISSING, "no group " + name); }
        // the particle count is the length of the step's first dataset (H5PartGetNumParticles)
        f->n       = 0;
        auto fields = names(f->step, 0);
        if (!fields.empty())
        {
            Handle d{H5Dopen2(f->step, fields[0].c_str(), H5P_DEFAULT), H5Dclose};
            Handle s{H5Dget_space(d.id), H5Sclose};
            f->n = uint64_t(H5Sget_simple_extent_npoints(s.id));
        }
        if (n) { *n = f->n; }
    });
}

int sphx_h5_write_field(sphx_h5file_t* f, const char* name, int type, const void* data)
{
    return guarded([&] {
        handle(f);
        if (f->mode == SPHX_H5_READ) { fail(SPHX_H5_ERR_ARG, "file opened for reading"); }
        hid_t g = stepGroup(f);
        if (name == nullptr || (data == nullptr && f->n > 0)) { fail(SPHX_H5_ERR_ARG, "null argument"); }
        const hid_t t = nativeType(type);
        const hsize_t dims = f->n;
        Handle        space{H5Screate_simple(1, &dims, nullptr), H5Sclose};
        // an existing dataset of the step is written over, as H5PartWriteData* does (H5Part.c:950-960)
        const std::string dn = datasetName(name);
        hid_t d = H5Lexists(g, dn.c_str(), H5P_DEFAULT) > 0
                      ? H5Dopen2(g, dn.c_str(), H5P_DEFAULT)
                      : H5Dcreate2(g, dn.c_str(), t, space.id, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (d < 0) { fail(SPHX_H5_ERR_IO, "cannot create dataset " + std::string(name)); }
        Handle dh{d, H5Dclose};
        Handle ds{H5Dget_space(d), H5Sclose};
        if (uint64_t(H5Sget_simple_extent_npoints(ds.id)) != f->n)
            fail(SPHX_H5_ERR_ARG, "dataset " + std::string(name) + " exists with another particle count");
        if (f->n == 0) { return; }
        const size_t es = typeSize(type);
        if (!onDevice(data))
        {
            check(H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "write " + std::string(name));
            return;
        }
        if (pinnedStaging(f, std::min<size_t>(kChunkBytes, f->n * es) + es))
        {
            pipelined(f, d, t, es, f->n, const_cast<void*>(data), true, name);
            return;
        }
        std::vector<char> host(f->n * es); // no page-locked memory: one pageable copy
        if (hipMemcpy(host.data(), data, host.size(), hipMemcpyDeviceToHost) != hipSuccess)
            fail(SPHX_H5_ERR_IO, "device to host copy of field " + std::string(name));
        check(H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, host.data()), "write " + std::string(name));
    });
}

int sphx_h5_read_field(sphx_h5file_t* f, const char* name, int type, void* data)
{
    return guarded([&] {
        handle(f);
        hid_t g = stepGroup(f);
        if (name == nullptr || (data == nullptr && f->n > 0)) { fail(SPHX_H5_ERR_ARG, "null argument"); }
        const hid_t t = nativeType(type);
        const std::string dn = datasetName(name);
        if (H5Lexists(g, dn.c_str(), H5P_DEFAULT) <= 0)
            fail(SPHX_H5_ERR_MISSING, "Could not read field: " + std::string(name));
        Handle d{H5Dopen2(g, dn.c_str(), H5P_DEFAULT), H5Dclose};
        Handle s{H5Dget_space(d.id), H5Sclose};
        if (uint64_t(H5Sget_simple_extent_npoints(s.id)) != f->n)
            fail(SPHX_H5_ERR_ARG, "field " + std::string(name) + " does not hold the step's particle count");
        if (f->n == 0) { return; }
        const size_t es = typeSize(type);
        if (!onDevice(data))
        {
            check(H5Dread(d.id, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, data), "read " + std::string(name));
            return;
        }
        if (pinnedStaging(f, std::min<size_t>(kChunkBytes, f->n * es) + es))
        {
            pipelined(f, d.id, t, es, f->n, data, false, name);
            return;
        }
        std::vector<char> host(f->n * es); // no page-locked memory: one pageable copy
        check(H5Dread(d.id, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, host.data()), "read " + std::string(name));
        if (hipMemcpy(data, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess)
            fail(SPHX_H5_ERR_IO, "host to device copy of field " + std::string(name));
    });
}

int sphx_h5_write_attrib(sphx_h5file_t* f, int step, const char* name, int type, const void* value, uint64_t count)
{
    return guarded([&] {
        handle(f);
        if (f->mode == SPHX_H5_READ) { fail(SPHX_H5_ERR_ARG, "file opened for reading"); }
        if (name == nullptr || value == nullptr || count == 0) { fail(SPHX_H5_ERR_ARG, "null or empty attribute"); }
        hid_t         owner = attribOwner(f, step);
        const hid_t   t     = nativeType(type);
        const hsize_t dims  = count;
        if (H5Aexists(owner, name) > 0) { check(H5Adelete(owner, name), "replace attribute"); }
        Handle space{H5Screate_simple(1, &dims, nullptr), H5Sclose};
        Handle a{H5Acreate2(owner, name, t, space.id, H5P_DEFAULT, H5P_DEFAULT), H5Aclose};
        if (a.id < 0) { fail(SPHX_H5_ERR_IO, "cannot create attribute " + std::string(name)); }
        check(H5Awrite(a.id, t, value), "write attribute " + std::string(name));
    });
}

int sphx_h5_attrib_info(sphx_h5file_t* f, int step, const char* name, int* type, uint64_t* count)
{
    return guarded([&] {
        handle(f);
        hid_t owner = attribOwner(f, step);
        if (name == nullptr || H5Aexists(owner, name) <= 0)
            fail(SPHX_H5_ERR_MISSING, "Attribute " + std::string(name ? name : "") + " does not exist");
        Handle a{H5Aopen(owner, name, H5P_DEFAULT), H5Aclose};
        Handle t{H5Aget_type(a.id), H5Tclose};
        Handle s{H5Aget_space(a.id), H5Sclose};
        if (type) { *type = typeOf(t.id); }
        if (count) { *count = uint64_t(H5Sget_simple_extent_npoints(s.id)); }
    });
}

int sphx_h5_read_attrib(sphx_h5file_t* f, int step, const char* name, int type, void* value, uint64_t count)
{
    int      fileType  = 0;
    uint64_t fileCount = 0;
    int      r         = sphx_h5_attrib_info(f, step, name, &fileType, &fileCount);
    if (r != SPHX_H5_OK) { return r; }
    return guarded([&] {
        nativeType(type);
        if (value == nullptr) { fail(SPHX_H5_ERR_ARG, "null argument"); }
        if (fileCount != count)
        {
            fail(SPHX_H5_ERR_TYPE, "Attribute " + std::string(name) + " size mismatch: in file " +
                                       std::to_string(fileCount) + ", but provided buffer has size " +
                                       std::to_string(count));
        }
        if (!convertible(fileType, type))
            fail(SPHX_H5_ERR_TYPE, "Reading attribute " + std::string(name) + " failed: type in file cannot convert");
        hid_t             owner = attribOwner(f, step);
        Handle            a{H5Aopen(owner, name, H5P_DEFAULT), H5Aclose};
        std::vector<char> buf(count * typeSize(fileType));
        check(H5Aread(a.id, nativeType(fileType), buf.data()), "read attribute " + std::string(name));
        convert(buf.data(), fileType, value, type, count);
    });
}

int sphx_h5_count(sphx_h5file_t* f, int step, int kind, int64_t* out)
{
    return guarded([&] {
        handle(f);
        if (out == nullptr || (kind == 0 && !step)) { fail(SPHX_H5_ERR_ARG, "fields exist per step only"); }
        *out = int64_t(names(kind == 0 ? stepGroup(f) : attribOwner(f, step), kind).size());
    });
}

int sphx_h5_name(sphx_h5file_t* f, int step, int kind, int64_t i, char* buf, size_t len)
{
    return guarded([&] {
        handle(f);
        if (buf == nullptr || len == 0 || (kind == 0 && !step)) { fail(SPHX_H5_ERR_ARG, "invalid argument"); }
        auto all = names(kind == 0 ? stepGroup(f) : attribOwner(f, step), kind);
        if (i < 0 || i >= int64_t(all.size())) { fail(SPHX_H5_ERR_MISSING, "no entry " + std::to_string(i)); }
        const std::string& name = all[size_t(i)];
        if (name.size() + 1 > len)
            fail(SPHX_H5_ERR_ARG, "name buffer too small: " + std::to_string(name.size() + 1) + " bytes needed");
        std::memcpy(buf, name.c_str(), name.size() + 1);
    });
}

} // extern "C"
