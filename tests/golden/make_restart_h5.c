/* Writes restart files with libhdf5 itself, using the call sequence of the
 * reference's RestartIO::write (src/io/restart.cpp:8-91: H5Fcreate with
 * default property lists; scalar root attributes time (double), iteration,
 * poly_order (int), gamma (double); 1-D datasets "solution" (double) then
 * "elem_n_sp" (int)) and of RestartIO::writeParallel without parallel HDF5
 * (src/io/restart.cpp:167-258: file <name>.rank<r>; attributes time,
 * iteration, num_procs, rank; dataset "solution").
 *
 *   make_restart_h5 <inputs.bin> <out.h5> <out_parallel_prefix> <rank> <num_procs>
 *
 * inputs.bin: int64 n_elem, int64 n_values, double time, int32 iteration,
 * int32 poly_order, double gamma, int32 elem_n_sp[n_elem], double
 * solution[n_values].  Built and run by make_restart_golden.py against an
 * installed libhdf5; the files it writes are committed as fixtures. */
#include <hdf5.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void die(const char* what) {
    fprintf(stderr, "make_restart_h5: %s\n", what);
    exit(1);
}

static void scalar_attr(hid_t loc, const char* name, hid_t type, const void* v) {
    hid_t sp = H5Screate(H5S_SCALAR);
    hid_t a = H5Acreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    if (a < 0 || H5Awrite(a, type, v) < 0) die(name);
    H5Aclose(a);
    H5Sclose(sp);
}

static void vector_dset(hid_t loc, const char* name, hid_t type, hsize_t n, const void* v) {
    hsize_t dims[1] = {n};
    hid_t sp = H5Screate_simple(1, dims, NULL);
    hid_t d = H5Dcreate2(loc, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (d < 0 || H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, v) < 0) die(name);
    H5Dclose(d);
    H5Sclose(sp);
}

int main(int argc, char** argv) {
    if (argc != 6) die("usage: make_restart_h5 <inputs.bin> <out.h5> <out_parallel_prefix> <rank> <num_procs>");
    FILE* f = fopen(argv[1], "rb");
    if (!f) die("cannot open inputs");
    int64_t n_elem, n_values;
    double time, gamma;
    int32_t iteration, poly_order;
    if (fread(&n_elem, 8, 1, f) != 1 || fread(&n_values, 8, 1, f) != 1 || fread(&time, 8, 1, f) != 1 ||
        fread(&iteration, 4, 1, f) != 1 || fread(&poly_order, 4, 1, f) != 1 || fread(&gamma, 8, 1, f) != 1)
        die("short header");
    int32_t* nsp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n_elem ? n_elem : 1));
    double* sol = (double*)malloc(sizeof(double) * (size_t)(n_values ? n_values : 1));
    if ((int64_t)fread(nsp, 4, (size_t)n_elem, f) != n_elem || (int64_t)fread(sol, 8, (size_t)n_values, f) != n_values)
        die("short data");
    fclose(f);

    hid_t file = H5Fcreate(argv[2], H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) die("H5Fcreate");
    int ival;
    scalar_attr(file, "time", H5T_NATIVE_DOUBLE, &time);
    ival = iteration;
    scalar_attr(file, "iteration", H5T_NATIVE_INT, &ival);
    ival = poly_order;
    scalar_attr(file, "poly_order", H5T_NATIVE_INT, &ival);
    scalar_attr(file, "gamma", H5T_NATIVE_DOUBLE, &gamma);
    vector_dset(file, "solution", H5T_NATIVE_DOUBLE, (hsize_t)n_values, sol);
    vector_dset(file, "elem_n_sp", H5T_NATIVE_INT, (hsize_t)n_elem, nsp);
    H5Fclose(file);

    int rank = atoi(argv[4]), nprocs = atoi(argv[5]);
    char path[4096];
    snprintf(path, sizeof path, "%s.rank%d", argv[3], rank);
    file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) die("H5Fcreate (parallel fallback)");
    scalar_attr(file, "time", H5T_NATIVE_DOUBLE, &time);
    ival = iteration;
    scalar_attr(file, "iteration", H5T_NATIVE_INT, &ival);
    scalar_attr(file, "num_procs", H5T_NATIVE_INT, &nprocs);
    scalar_attr(file, "rank", H5T_NATIVE_INT, &rank);
    vector_dset(file, "solution", H5T_NATIVE_DOUBLE, (hsize_t)n_values, sol);
    H5Fclose(file);
    free(nsp);
    free(sol);
    return 0;
}
