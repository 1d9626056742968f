// zhijian/io.hpp -- restart files and visualisation output (drop-in API).
//
// RestartIO: the reference's include/io/restart.hpp:12-70.  Files are HDF5
// (file format 0) with the reference's layout (src/io/restart.cpp:8-91):
// root attributes time, iteration, poly_order, gamma; datasets "solution"
// (all sol_pts, element by element, variable fastest) and "elem_n_sp".  They
// are written and read by the library's own HDF5 codec, so no HDF5 library is
// needed, and files written by libhdf5 read back the same.
//   * writeParallel/readParallel keep the reference's behaviour without
//     parallel HDF5 (restart.cpp:182-183, 276): one file "<name>.rank<r>" per rank
//     with attributes time, iteration, num_procs, rank and dataset "solution".
//     readParallel fills `solution` from the local mesh's element types and
//     params.poly_order (the reference leaves it empty, restart.cpp:318-319).
//   * every failure returns false with errorMessage() set (the reference does
//     not check HDF5 return codes while reading).
//
// VTKWriter: the reference's include/io/visualization.hpp:13-74 and
// src/io/visualization.cpp.  writeVTU/writeLegacy/writePVTU/writeParallelVTU
// produce the reference's text byte for byte by default: each element
// subdivided into a lattice of setSubdivisions() cells split into triangles,
// point coordinates from the element's shape functions, and the state of
// the element's FIRST solution point at every lattice point
// (visualization.cpp:327, 360).  Two switches the reference declares and
// ignores take effect here:
//   * setHighOrderOutput(true) (the default) with params.compat == Textbook
//     interpolates the solution polynomial to the lattice points instead;
//   * setBinary(true) writes VTU data arrays in appended raw binary
//     (header_type UInt64); legacy files stay ASCII.
//
// ResidualWriter: visualization.hpp:124-141 (CSV residual history).
//
// TecplotWriter and SurfaceDataWriter are declared by the reference
// (visualization.hpp:77-121) but never defined, so they are not provided.
#pragma once

#include "zhijian/fr_solver.hpp"
#include "zhijian/mesh.hpp"
#include "zhijian/types.hpp"

#include <array>
#include <fstream>
#include <string>
#include <vector>

namespace zhijian {

class RestartIO {
public:
    RestartIO() = default;
    bool write(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
               const SimParams& params, Real time, int iteration);
    bool read(const std::string& filename, Mesh& mesh, std::vector<ElementSolution>& solution,
              SimParams& params, Real& time, int& iteration);
    bool writeParallel(const std::string& filename, int rank, int num_procs, const Mesh& local_mesh,
                       const std::vector<ElementSolution>& solution, const SimParams& params, Real time,
                       int iteration);
    bool readParallel(const std::string& filename, int rank, int num_procs, Mesh& local_mesh,
                      std::vector<ElementSolution>& solution, SimParams& params, Real& time, int& iteration);
    const std::string& errorMessage() const { return error_msg_; }

private:
    std::string error_msg_;
};

class VTKWriter {
public:
    VTKWriter() = default;
    bool writeLegacy(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
                     const SimParams& params);
    bool writeVTU(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
                  const SimParams& params);
    bool writeParallelVTU(const std::string& base_filename, int rank, int num_procs, const Mesh& local_mesh,
                          const std::vector<ElementSolution>& solution, const SimParams& params);
    bool writePVTU(const std::string& filename, int num_procs, const std::vector<std::string>& piece_filenames);
    void setBinary(bool binary) { binary_ = binary; }
    void setHighOrderOutput(bool high_order) { high_order_ = high_order; }
    void setSubdivisions(int n) { subdivisions_ = n; }
    const std::string& errorMessage() const { return error_msg_; }

private:
    bool write(const std::string& filename, const Mesh& mesh, const std::vector<ElementSolution>& solution,
               const SimParams& params, int format);
    std::string error_msg_;
    bool binary_ = false;
    bool high_order_ = true;
    int subdivisions_ = 4;
};

class ResidualWriter {
public:
    ResidualWriter() = default;
    bool open(const std::string& filename);
    void write(int iteration, Real time, Real residual, const std::array<Real, N_VARS>& component_residuals);
    void close();

private:
    std::ofstream file_;
};

}  // namespace zhijian
