/*
 * spherocyl_files.h -- the reference's text formats as C entry points of
 * libspherocyls.so (host only: no GPU needed).  SpherocylBox writes every file
 * through these functions.
 */
#ifndef SPHEROCYL_FILES_H
#define SPHEROCYL_FILES_H

#ifdef __cplusplus
extern "C" {
#endif

/* save_positions (src/cpp/spherocyl_box.cpp:410-423): header line
 * "N L packing gamma gamma_total strain_rate step", then "x y phi R A" per
 * particle, C++ stream formatting at 14 significant digits.  Returns SC_OK, or
 * SC_ERR_ARG when the file cannot be written. */
int sc_write_positions(const char* path, int n, double L, double packing, double gamma,
                       double gamma_total, double strain_rate, double step, const double* x,
                       const double* y, const double* phi, const double* R, const double* A);

/* Reads a save_positions file (the restart read of src/cpp/main.cpp:130-157 via
 * DatFileInput, src/cpp/dat_file_input.cpp:36-58).  header[6] gets
 * L packing gamma gamma_total strain_rate step; *n the number of particle rows;
 * at most cap rows are stored (arrays may be NULL when cap is 0).  The GPU
 * code's sd*.dat files, whose header is "N Lx Ly packing ..." (the 8-field
 * case of src/cuda/main_2point_qscompress.cpp:189-201), are read too when
 * Lx == Ly; other header lengths are SC_ERR_ARG. */
int sc_read_positions(const char* path, int cap, int* n, double* header, double* x, double* y,
                      double* phi, double* R, double* A);

/* The GPU code's save_positions (src/cuda/strain_dynamics.cu:820-843), file
 * sd%010lu.dat of its quasi-static protocols: header
 * "N Lx Ly packing gamma gamma_total strain_rate step" printed
 * "%d %.17g %.17g %f %.17g %.14g %g %g", then "x y phi R A" per particle
 * printed "%.17g %.17g %.17g %.2g %.17g" (R at two digits, as there).
 * Returns SC_OK, or SC_ERR_ARG when the file cannot be written. */
int sc_write_positions_sd(const char* path, int n, double Lx, double Ly, double packing,
                          double gamma, double gamma_total, double strain_rate, double step,
                          const double* x, const double* y, const double* phi, const double* R,
                          const double* A);

/* save_positions_bin (src/cuda/strain_dynamics.cu:845-875), file sd%010lu.bin:
 * the particle count as a native int, then native doubles Lx Ly packing gamma
 * gamma_total strain_rate step, then x y phi R A of each particle in turn. */
int sc_write_positions_bin(const char* path, int n, double Lx, double Ly, double packing,
                           double gamma, double gamma_total, double strain_rate, double step,
                           const double* x, const double* y, const double* phi, const double* R,
                           const double* A);

/* Reads a save_positions_bin file (the reference reads it field by field with
 * BinFileInput, src/cuda/bin_file_input.cpp:60-140).  header[7] gets
 * Lx Ly packing gamma gamma_total strain_rate step; *n the particle count; at
 * most cap particles are stored.  SC_ERR_ARG for a missing, short or
 * inconsistent file (its size must be 4 + 8 * (7 + 5 N) bytes). */
int sc_read_positions_bin(const char* path, int cap, int* n, double* header, double* x, double* y,
                          double* phi, double* R, double* A);

/* One stress/energy line "t v0 v1 ...\n" as the reference's SE writers format
 * it (src/cpp/strain_dynamics.cpp:291, resize_dynamics.cpp:298): C++ stream
 * output of each value at `precision` significant digits (6 is the stream
 * default the main protocols use, 14 the run_strain(L) / resize_box(phi)
 * overloads).  Writes at most cap bytes including the terminating NUL; returns
 * the line length, or -1 if it does not fit. */
int sc_format_se_line(char* buf, int cap, unsigned long t, int nvals, const double* vals,
                      int precision);

#ifdef __cplusplus
}
#endif

#endif
