/* matlab_calculate_return.h -- declaration MATLAB's loadlibrary reads for
 * libConnect.so (the connector library built from
 * matrix-permanent-calculator_amd/host/lib_connect.cpp).  Same name and
 * argument list as the reference connector's header
 * (matlab_calculate_return.h, py_connector.c:152): the dense int matrix
 * `mat` (nov x nov, row-major), the algorithm ID (4-8 exact, on the GPU),
 * threads and approximation parameters (accepted, unused), nov and nnz. */
#ifndef MATLAB_CALCULATE_RETURN_H
#define MATLAB_CALCULATE_RETURN_H
#ifdef __cplusplus
extern "C" {
#endif
double matlab_calculate_return(int* mat, int algorithm, int nt, int x, int y, int z, int nov,
                               int nnz);
#ifdef __cplusplus
}
#endif
#endif
