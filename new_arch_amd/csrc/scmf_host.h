// scmf_host.h — host-side formats of the SCMF engine: configuration.dat and
// polymer.dat readers, density_<step>.dat / coord.dat writers, VTK converter.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/scmf.h"

namespace scmf {

// Owning storage behind a scmf_desc.
struct HostDesc {
  int32_t T = 0;
  std::vector<double> xn;
  double L[3] = {0, 0, 0};
  int32_t n[3] = {0, 0, 0};
  double dtNxi = 0, xiN = 0, kT = 0;
  int32_t time = 0;
  int32_t nref = 0;
  std::vector<int32_t> poly_length;
  std::vector<int64_t> n_polymers;
  std::vector<int32_t> monotype;
  std::vector<uint8_t> connection;
  scmf_desc view() const;
};

// Read configuration.dat (init_config.cu:487-514).  Returns "" or an error.
std::string read_configuration(const char* path, HostDesc* d);
// Read polymer.dat (init_config.cu:210-343).  Returns "" or an error.
std::string read_polymer(const char* path, HostDesc* d);

// Validate a description; returns "" or the first problem found.
std::string validate_desc(const scmf_desc* d);

// density_<step>.dat name exactly as io.cpp:16-19 builds it.
std::string density_file_name(int32_t step);

// Write density_<step>.dat (io.cpp:27-48).
std::string write_density_file(const char* path, const int32_t n[3], const double L[3], int T,
                               int64_t n_cells, const uint32_t* counts, const float* omega,
                               double field_scaling);

// The VTK file density_to_vtk(density_name, path) makes of the density file
// write_density_file writes from the same data.
std::string write_vtk_file(const char* path, const char* density_name, const int32_t n[3],
                           const double L[3], int T, int64_t n_cells, const uint32_t* counts,
                           const float* omega, double field_scaling);

// Write coord.dat (io.cpp:101-107).
std::string write_coord_file(const char* path, const float* xyz, int64_t n_beads);

// vtk.c:42-76 generalised to 2T columns per row.
std::string density_to_vtk(const char* in, const char* out);

}  // namespace scmf
