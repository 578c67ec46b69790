// Material models behind the reference's MaterialModel interface and factory
// names (reference src/energies/material_model.h:12-119,
// src/factories/material_model_factory.cpp:12-49).
//
// A model carries its MaterialConfig and a device type id; per-element
// evaluation happens inside the element kernels. The batched entry points
// evaluate the model on many points on the GPU (the reference evaluates one
// point per call on the CPU).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <string>

#include "config.h"
#include "factory.h"

namespace mfem {

class MaterialModel {
 public:
  static std::string name() { return "base"; }
  explicit MaterialModel(const std::shared_ptr<MaterialConfig>& config) : config_(config) {}
  virtual ~MaterialModel() = default;

  const std::shared_ptr<MaterialConfig>& config() const { return config_; }
  MaterialModelType type() const { return config_->material_model; }

  // Batched evaluation on the 6-vector s (device pointers, n points,
  // column-major 6x6 Hessians). psd_fix mirrors hessian(s, psd_fix).
  void energy_gradient_hessian(int n, const double* s, bool psd_fix, double* E, double* g,
                               double* H, hipStream_t st) const;

 protected:
  std::shared_ptr<MaterialConfig> config_;
};

#define MFEM_MATERIAL(Cls, Name)                                                   \
  class Cls : public MaterialModel {                                               \
   public:                                                                         \
    static std::string name() { return Name; }                                     \
    explicit Cls(const std::shared_ptr<MaterialConfig>& c) : MaterialModel(c) {} \
  };
MFEM_MATERIAL(StableNeohookean, "Stable-Neohookean")
MFEM_MATERIAL(Neohookean, "Neohookean")
MFEM_MATERIAL(Corotational, "Corot")
MFEM_MATERIAL(FixedCorotational, "FixedCorot")
MFEM_MATERIAL(ArapModel, "ARAP")
MFEM_MATERIAL(Fung, "Fung")
#undef MFEM_MATERIAL

class MaterialModelFactory
    : public Factory<MaterialModelType, MaterialModel, std::shared_ptr<MaterialConfig>> {
 public:
  MaterialModelFactory();
};

// True for the models whose element kernels are implemented on the device.
bool material_supported_on_device(MaterialModelType t);

}  // namespace mfem
