/* Surface deposition counts and settled positions as HDF5: the layouts of
 * TriangleMeshHdf5Manager (TriangleMeshHdf5Manager.cpp:41-178) and
 * SettledParticleHDF5Writer (SettledParticleHDF5Writer.cpp:4-114), written and
 * read with the HDF5 C library.  Host only; C so that the library pulls in
 * nothing but libhdf5. */
#include "start_hdf5.h"

#include <hdf5.h>
#include <inttypes.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <unistd.h>

static _Thread_local char g_err[512];

static int fail(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
    return -1;
}

const char* start_hdf5_last_error(void) { return g_err; }

double start_triangle_area(const double* t)
{
    const double ax = t[3] - t[0], ay = t[4] - t[1], az = t[5] - t[2];
    const double bx = t[6] - t[0], by = t[7] - t[1], bz = t[8] - t[2];
    const double cx = ay * bz - az * by, cy = az * bx - ax * bz, cz = ax * by - ay * bx;
    return sqrt((cx * cx + cy * cy + cz * cz) / 4.0);
}

/* size_t as the reference stores it (TriangleMeshHdf5Manager.cpp:119-131) */
static hid_t counter_type(void)
{
    return sizeof(size_t) == sizeof(unsigned long) ? H5T_NATIVE_ULONG : H5T_NATIVE_ULLONG;
}

/* one 1-D dataset of `dims` elements in group g */
static int write_dataset(hid_t g, const char* name, hid_t type, const void* data, hsize_t dims)
{
    hid_t space = H5Screate_simple(1, &dims, NULL);
    if (space < 0) return -1;
    hid_t ds = H5Dcreate2(g, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = ds < 0 ? -1 : 0;
    if (rc == 0 && H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0) rc = -1;
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(space);
    return rc;
}

int start_write_triangle_hdf5(const char* path, int64_t n_tris, const int64_t* ids, const double* nodes,
                              const int32_t* tris, const uint64_t* counts)
{
    if (!path) return fail("null path");
    if (n_tris < 0) return fail("negative triangle count");
    if (n_tris > 0 && (!ids || !nodes || !tris || !counts)) return fail("null triangle array");
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);  /* errors come back as return codes, not on stderr */
    if (access(path, F_OK) == 0 && unlink(path) != 0) return fail("cannot remove existing file %s", path);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("failed to create HDF5 file: %s", path);
    int rc = 0;
    for (int64_t k = 0; k < n_tris && rc == 0; ++k) {
        char name[64];
        snprintf(name, sizeof name, "Triangle_%" PRId64, ids[k]);
        hid_t g = H5Gcreate2(f, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (g < 0) {
            rc = fail("failed to create group %s (duplicate id?)", name);
            break;
        }
        double xyz[9];
        for (int v = 0; v < 3; ++v)
            for (int c = 0; c < 3; ++c) xyz[3 * v + c] = nodes[3 * (int64_t)tris[3 * k + v] + c];
        const double area = start_triangle_area(xyz);
        const size_t count = (size_t)counts[k];
        if (write_dataset(g, "Coordinates", H5T_NATIVE_DOUBLE, xyz, 9) < 0 ||
            write_dataset(g, "Area", H5T_NATIVE_DOUBLE, &area, 1) < 0 ||
            write_dataset(g, "Counter", counter_type(), &count, 1) < 0)
            rc = fail("failed to write the datasets of %s", name);
        H5Gclose(g);
    }
    if (H5Fclose(f) < 0 && rc == 0) rc = fail("failed to close %s", path);
    return rc;
}

static int read_dataset(hid_t f, const char* group, const char* name, hid_t type, void* data, hsize_t want)
{
    hid_t g = H5Gopen2(f, group, H5P_DEFAULT);
    if (g < 0) return fail("no group %s", group);
    hid_t ds = H5Dopen2(g, name, H5P_DEFAULT);
    int rc = 0;
    if (ds < 0) {
        rc = fail("no dataset %s in group %s", name, group);
    } else {
        hid_t space = H5Dget_space(ds);
        hsize_t dims = 0;
        if (space < 0 || H5Sget_simple_extent_ndims(space) != 1 || H5Sget_simple_extent_dims(space, &dims, NULL) < 0 ||
            dims != want)
            rc = fail("dataset %s/%s is not one-dimensional with %llu elements", group, name, (unsigned long long)want);
        else if (H5Dread(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0)
            rc = fail("failed to read %s/%s", group, name);
        if (space >= 0) H5Sclose(space);
        H5Dclose(ds);
    }
    H5Gclose(g);
    return rc;
}

int start_read_triangle_hdf5(const char* path, int64_t first_id, int64_t* n, double* coords, double* areas,
                             uint64_t* counts)
{
    if (!path || !n) return fail("null argument");
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("failed to open HDF5 file: %s", path);
    H5G_info_t info;
    if (H5Gget_info(f, &info) < 0) {
        H5Fclose(f);
        return fail("failed to list the groups of %s", path);
    }
    const int64_t cap = *n, num = (int64_t)info.nlinks;
    *n = num;
    int rc = 0;
    if (cap > 0) {
        if (cap < num) rc = fail("capacity %" PRId64 " below the %" PRId64 " groups of the file", cap, num);
        else if (!coords || !areas || !counts) rc = fail("null output array");
    }
    for (int64_t k = 0; cap > 0 && rc == 0 && k < num; ++k) {
        char name[64];
        snprintf(name, sizeof name, "Triangle_%" PRId64, first_id + k);
        size_t count = 0;
        if (read_dataset(f, name, "Coordinates", H5T_NATIVE_DOUBLE, coords + 9 * k, 9) < 0 ||
            read_dataset(f, name, "Area", H5T_NATIVE_DOUBLE, areas + k, 1) < 0 ||
            read_dataset(f, name, "Counter", counter_type(), &count, 1) < 0)
            rc = -1;
        counts[k] = (uint64_t)count;
    }
    H5Fclose(f);
    return rc;
}

int start_write_settled_hdf5(const char* path, int64_t n, const double* xyz, const char* dataset)
{
    if (!path) return fail("null path");
    if (n < 0) return fail("negative particle count");
    if (n == 0) {
        fail("no particles to save: no file written");
        return 1;
    }
    if (!xyz) return fail("null position array");
    H5Eset_auto2(H5E_DEFAULT, NULL, NULL);
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("failed to create HDF5 file: %s", path);
    const hsize_t dims[2] = {(hsize_t)n, 3};
    hid_t space = H5Screate_simple(2, dims, NULL);
    hid_t ds = space < 0 ? -1
                         : H5Dcreate2(f, dataset ? dataset : "settled_particles", H5T_NATIVE_DOUBLE, space,
                                      H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    int rc = 0;
    if (ds < 0) rc = fail("failed to create the dataset in %s", path);
    else if (H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, xyz) < 0)
        rc = fail("failed to write %s", path);
    if (ds >= 0) H5Dclose(ds);
    if (space >= 0) H5Sclose(space);
    if (H5Fclose(f) < 0 && rc == 0) rc = fail("failed to close %s", path);
    return rc;
}
