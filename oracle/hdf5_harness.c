/* TEST INFRASTRUCTURE ONLY (never linked into the product): support for the
 * reference drivers' HDF5 readers, which oracle/Makefile compiles from their
 * own sources under oracle/_ref -- the APS driver's read_data
 * (reconstruct/real_data/APS_Data_Format/XT_HDFIO.c:11-190) and the standard
 * driver's read_ProjWeightData / read_AngleTimeReconList
 * (reconstruct/real_data/Standard_Data_Format/XT_HDFIO.c:11-215).  Those
 * readers ask MPI for the rank count, so h5x_mpi_up() initialises MPI once
 * (one rank).  h5x_write() writes the small fixture files the tests read
 * with both the reference readers and the product's.
 */
#include <hdf5.h>
#include <mpi.h>
#include <stdint.h>
#include <string.h>

int h5x_mpi_up(void) {
  int up = 0;
  MPI_Initialized(&up);
  if (!up) {
    int argc = 0;
    char** argv = NULL;
    if (MPI_Init(&argc, &argv) != MPI_SUCCESS) return -1;
  }
  return 0;
}

/* Writes dataset `name` (an absolute path; its parent groups are created)
 * of `rank` dimensions into file `path`: a new file when create != 0, else
 * an existing one.  kind 0: uint16, 1: float32.  Returns 0 or -1. */
int h5x_write(const char* path, int create, const char* name, int rank,
              const int64_t* dims, int kind, const void* data) {
  hsize_t d[8];
  if (rank < 1 || rank > 8) return -1;
  for (int i = 0; i < rank; ++i) d[i] = (hsize_t)dims[i];
  hid_t f = create ? H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
                   : H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT);
  if (f < 0) return -1;
  hid_t lcpl = H5Pcreate(H5P_LINK_CREATE);
  H5Pset_create_intermediate_group(lcpl, 1);
  hid_t type = kind == 0 ? H5T_NATIVE_UINT16 : H5T_NATIVE_FLOAT;
  hid_t space = H5Screate_simple(rank, d, NULL);
  hid_t ds = H5Dcreate2(f, name, type, space, lcpl, H5P_DEFAULT, H5P_DEFAULT);
  int rc = -1;
  if (ds >= 0) {
    rc = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) < 0 ? -1 : 0;
    H5Dclose(ds);
  }
  H5Sclose(space);
  H5Pclose(lcpl);
  H5Fclose(f);
  return rc;
}
