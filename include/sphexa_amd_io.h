/* sphexa_amd I/O: particle fields and attributes in the H5Part layout of HDF5 files, the format the
 * reference writes and reads (main/src/io/ifile_io_hdf5.cpp through extern/h5part), so that glass
 * initial conditions, outputs and checkpoints pass between the two.
 *
 * Layout (extern/h5part/H5Part.c): the root group holds one group per step, "Step#0", "Step#1", ...;
 * a step group holds one 1-D dataset per field (all of the step's particle count, native HDF5
 * types: double, float, int32, int64, char) and the step attributes (1-D attributes of the group);
 * file attributes sit on the root group.
 *
 * One rank's share of a file: the library opens files serially, as the reference's H5Part build
 * without parallel HDF5 does on one rank (h5part_wrapper.hpp:332-354). Field pointers may be device
 * (HIP) or host memory. A device field is read or written after a device synchronisation, in chunks
 * through page-locked host buffers the handle keeps until it closes; the call returns when the
 * data is in the file or in device memory. Calls on one handle must not overlap. Separate library (libsphexa_amd_io.so, linked against HDF5) so that the compute library
 * does not depend on HDF5. */
#ifndef SPHEXA_AMD_IO_H
#define SPHEXA_AMD_IO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* open modes (H5PART_READ / H5PART_WRITE / H5PART_APPEND, H5Part.h:42-44) */
#define SPHX_H5_READ 1
#define SPHX_H5_WRITE 2
#define SPHX_H5_APPEND 4

/* element types of fields and attributes: the reference's IO::Types (main/src/io/ifile_io.hpp:42-47)
 * as H5Part stores them (h5part_wrapper.hpp:56-106): unsigned as int32, uint64 as int64, char and
 * uint8 as char */
typedef enum sphx_h5_type
{
    SPHX_H5_F64 = 0,
    SPHX_H5_F32 = 1,
    SPHX_H5_I32 = 2, /* int, unsigned */
    SPHX_H5_I64 = 3, /* int64_t, uint64_t */
    SPHX_H5_I8  = 4  /* char, uint8_t */
} sphx_h5_type;

/* status: 0 on success, else an error whose message sphx_h5_last_error() returns (per thread) */
#define SPHX_H5_OK 0
#define SPHX_H5_ERR_ARG -1     /* invalid argument or call out of order */
#define SPHX_H5_ERR_IO -2      /* an HDF5 or HIP call failed */
#define SPHX_H5_ERR_MISSING -3 /* no such field, attribute or step (std::out_of_range in the reference) */
#define SPHX_H5_ERR_TYPE -4    /* attribute size or type that cannot be converted safely */

typedef struct sphx_h5file sphx_h5file_t;

const char* sphx_h5_last_error(void);

/* Open `path` (openH5Part, h5part_wrapper.hpp:332-354): SPHX_H5_READ an existing file,
 * SPHX_H5_WRITE a new (truncated) file, SPHX_H5_APPEND an existing file for more steps. */
int sphx_h5_open(const char* path, int mode, sphx_h5file_t** out);
int sphx_h5_close(sphx_h5file_t* f);

/* Steps in the file (H5PartGetNumSteps) */
int sphx_h5_num_steps(sphx_h5file_t* f, int64_t* out);

/* Writer: start the next step, "Step#<number of steps>", with n particles per field
 * (H5PartWriter::addStep, ifile_io_hdf5.cpp:55-77). */
int sphx_h5_add_step(sphx_h5file_t* f, uint64_t n);

/* Reader: make `step` current (negative: the last step), H5PartReader::setStep
 * (ifile_io_hdf5.cpp:180-210) in independent mode; *n receives its particle count (the length of
 * its first dataset, H5PartGetNumParticles). */
int sphx_h5_set_step(sphx_h5file_t* f, int64_t step, uint64_t* n);

/* Field of the current step: n elements of `type` from/to `data` (device or host). Reading converts
 * from the type in the file, as H5PartReadData* does through HDF5. */
int sphx_h5_write_field(sphx_h5file_t* f, const char* name, int type, const void* data);
int sphx_h5_read_field(sphx_h5file_t* f, const char* name, int type, void* data);

/* Attributes of the current step (step = 1) or of the file (step = 0), `count` elements of `type`
 * (host memory). Reading follows readAttribute (h5part_wrapper.hpp:183-222): the count must match
 * the file's, and the values convert when both types are floating point or the requested type is
 * the common type of both; otherwise SPHX_H5_ERR_TYPE. */
int sphx_h5_write_attrib(sphx_h5file_t* f, int step, const char* name, int type, const void* value,
                         uint64_t count);
int sphx_h5_read_attrib(sphx_h5file_t* f, int step, const char* name, int type, void* value, uint64_t count);

/* Names: number of fields of the current step (step = 1 with kind 0), of its attributes (step = 1,
 * kind 1) or of the file attributes (step = 0, kind 1), and name i of them into buf (a name that does
 * not fit in len bytes with its terminator is SPHX_H5_ERR_ARG, buf untouched: call again with a
 * larger buffer); the size and element type of an attribute. */
int sphx_h5_count(sphx_h5file_t* f, int step, int kind, int64_t* out);
int sphx_h5_name(sphx_h5file_t* f, int step, int kind, int64_t i, char* buf, size_t len);
int sphx_h5_attrib_info(sphx_h5file_t* f, int step, const char* name, int* type, uint64_t* count);

#ifdef __cplusplus
}
#endif

#endif
