/*
 * Reads a checkpoint in the reference's layout (Output/h5out.c: Grid/T,
 * Grid/Np[Nz][Nr], Data/Cells[Ntot][NUM_Q + NUM_FACES + 1] = prim, Phi,
 * piph per cell) and writes T, the two dimensions of Np and its values
 * (int64), then the two dimensions of Cells and its values (doubles) to a
 * raw file, for the binding test.  Test infrastructure only.
 * Usage: h5cells <checkpoint.h5> <out.bin>
 */
#include <hdf5.h>
#include <stdio.h>
#include <stdlib.h>

int main(int argc, char **argv)
{
   if (argc != 3) return 2;
   hid_t f = H5Fopen(argv[1], H5F_ACC_RDONLY, H5P_DEFAULT);
   if (f < 0) return 1;
   double T = 0.0;
   hid_t d = H5Dopen2(f, "/Grid/T", H5P_DEFAULT);
   H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &T);
   H5Dclose(d);
   d = H5Dopen2(f, "/Grid/Np", H5P_DEFAULT);
   hid_t s = H5Dget_space(d);
   hsize_t ndims[2] = {0, 0};
   H5Sget_simple_extent_dims(s, ndims, NULL);
   int *np = (int *)malloc(sizeof(int) * ndims[0] * ndims[1]);
   H5Dread(d, H5T_NATIVE_INT, H5S_ALL, H5S_ALL, H5P_DEFAULT, np);
   H5Sclose(s);
   H5Dclose(d);
   d = H5Dopen2(f, "/Data/Cells", H5P_DEFAULT);
   s = H5Dget_space(d);
   hsize_t dims[2] = {0, 0};
   H5Sget_simple_extent_dims(s, dims, NULL);
   double *buf = (double *)malloc(sizeof(double) * dims[0] * dims[1]);
   H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf);
   H5Sclose(s);
   H5Dclose(d);
   H5Fclose(f);
   FILE *o = fopen(argv[2], "wb");
   if (!o) return 1;
   long n[2] = {(long)dims[0], (long)dims[1]}, nn[2] = {(long)ndims[0], (long)ndims[1]};
   fwrite(&T, sizeof(double), 1, o);
   fwrite(nn, sizeof(long), 2, o);
   for (hsize_t i = 0; i < ndims[0] * ndims[1]; ++i) {
      long v = np[i];
      fwrite(&v, sizeof(long), 1, o);
   }
   fwrite(n, sizeof(long), 2, o);
   fwrite(buf, sizeof(double), dims[0] * dims[1], o);
   fclose(o);
   free(buf);
   free(np);
   return 0;
}
