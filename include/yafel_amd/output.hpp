// YAFEL's output layer over the C ABI, with the reference's names:
//
//   OutputData(data, name[, location, type, dof_mask])  include/output/OutputData.hpp:19-57
//   OutputFrame(...).addData(...)                        include/output/OutputFrame.hpp:17-38
//   SimulationOutput(fname_base, BackendType::VTU)       include/output/SimulationOutput.hpp:25-101
//     .captureFrame(fesys[, captureFunction])            (default: Solution_<comp> scalars)
//
// The frame is written by yafel_hip_vtu_write, which restates VTUBackend and
// OutputMesh.  Two deliberate differences: the reference appends every frame
// to one fname_base.vtu whose closing tags are written once (a second frame
// gives malformed XML), here frame k > 0 goes to fname_base_<k>.vtu; and the
// HDF5 backend, which the reference builds only with USE_HDF5, is not part of
// this build (std::invalid_argument).
#pragma once

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "yafel_amd/fe.hpp"

namespace yafel {

enum class BackendType { VTU, HDF5 };

class OutputData {
 public:
  enum class DataType : int { Scalar, Vector, Tensor };
  enum class DataLocation { Point, Cell };

  OutputData(Vector<double> &d, const std::string &name)
      : OutputData(d, name, DataLocation::Point, DataType::Scalar, {1}) {}
  OutputData(Vector<double> &d, const std::string &name, DataLocation dL, DataType dT, const std::vector<int> &mask)
      : data(d.data()), size(d.size()), dof_mask(mask), name(name), dataType(dT), dataLocation(dL) {}
  OutputData(std::vector<double> &d, const std::string &name, DataLocation dL, DataType dT,
             const std::vector<int> &mask)
      : data(d.data()), size(d.size()), dof_mask(mask), name(name), dataType(dT), dataLocation(dL) {}

  const double *data;
  std::size_t size;
  std::vector<int> dof_mask;
  std::string name;
  DataType dataType;
  DataLocation dataLocation;
};

class OutputFrame {
 public:
  explicit OutputFrame(const DoFManager &dofm) : dofm(&dofm) {}
  const DoFManager *dofm;
  std::vector<std::shared_ptr<OutputData>> point_data;
  std::vector<std::shared_ptr<OutputData>> cell_data;
  double time{0};
  void addData(std::shared_ptr<OutputData> data) {
    if (data->dataLocation == OutputData::DataLocation::Point) point_data.push_back(data);
    if (data->dataLocation == OutputData::DataLocation::Cell) cell_data.push_back(data);
  }
};

class SimulationOutput {
 public:
  explicit SimulationOutput(std::string fname_base = "output", BackendType backendType = BackendType::VTU)
      : fname_base_(std::move(fname_base)) {
    if (backendType != BackendType::VTU)
      throw std::invalid_argument("SimulationOutput: only the VTU backend is part of this build");
  }

  template <typename T>
  void captureFrame(T &fes, const std::function<void(T &, OutputFrame &)> &captureFunction = default_capture<T>()) {
    const DoFManager &dofm = fes.getDoFManager();
    OutputFrame frame(dofm);
    captureFunction(fes, frame);
    write(frame);
  }

  // SimulationOutput.hpp:58-92: every solution component as scalar data
  // "Solution_<i>", cell data when there is one value per cell and component
  template <typename T>
  static std::function<void(T &, OutputFrame &)> default_capture() {
    return [](T &fes, OutputFrame &frame) {
      frame.time = fes.currentTime();
      auto &dofm = fes.getDoFManager();
      const int dofpn = dofm.dof_per_node;
      auto &solution = fes.getSolution();
      std::vector<int> dof_mask(dofpn, 0);
      auto loc = OutputData::DataLocation::Point;
      if ((std::size_t)dofpn * dofm.nCells() == solution.size()) loc = OutputData::DataLocation::Cell;
      for (int i = 0; i < dofpn; ++i) {
        dof_mask[i] = 1;
        frame.addData(std::make_shared<OutputData>(solution, "Solution_" + std::to_string(i), loc,
                                                   OutputData::DataType::Scalar, dof_mask));
        dof_mask[i] = 0;
      }
    };
  }

 private:
  void write(const OutputFrame &frame) {
    const DoFManager &dofm = *frame.dofm;
    std::vector<double> xyz(dofm.dof_nodes.size() * 3);
    for (std::size_t i = 0; i < dofm.dof_nodes.size(); ++i)
      for (int d = 0; d < 3; ++d) xyz[3 * i + d] = dofm.dof_nodes[i](d);
    std::vector<int32_t> cells(dofm.elements.begin(), dofm.elements.end());
    std::vector<yafel_hip_output_data> recs;
    for (const auto *list : {&frame.point_data, &frame.cell_data})
      for (const auto &d : *list) {
        yafel_hip_output_data r{};
        r.name = d->name.c_str();
        r.data = d->data;
        r.on_device = 0;
        r.dof_per_loc = (int)d->dof_mask.size();
        for (std::size_t c = 0; c < d->dof_mask.size() && c < 32; ++c)
          if (d->dof_mask[c]) r.mask |= 1u << c;
        r.data_type = (int)d->dataType;
        r.location = d->dataLocation == OutputData::DataLocation::Point ? 0 : 1;
        const std::size_t nloc = r.location == 0 ? dofm.dof_nodes.size() : (std::size_t)dofm.nCells();
        if (d->size != nloc * d->dof_mask.size())
          throw std::runtime_error("VTUBackend::write_data: Incompatible dof_mask length. Data length must be a "
                                   "multiple of dof_mask length: data = " + d->name);
        recs.push_back(r);
      }
    const std::string base = frames_ == 0 ? fname_base_ : fname_base_ + "_" + std::to_string(frames_);
    hipdetail::check(yafel_hip_vtu_write(base.c_str(), dofm.nsd(), dofm.polyOrder, (int64_t)dofm.dof_nodes.size(),
                                         xyz.data(), (int64_t)dofm.nCells(), cells.data(), (int)recs.size(),
                                         recs.data()));
    ++frames_;
  }

  std::string fname_base_;
  int frames_ = 0;
};

}  // namespace yafel
