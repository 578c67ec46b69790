#include "material_model.h"

#include <vector>

#include "../kernels.h"
#include "device.h"

namespace mfem {

void MaterialModel::energy_gradient_hessian(int n, const double* s, bool psd_fix, double* E,
                                            double* g, double* H, hipStream_t st) const {
  std::vector<int> t(n, (int)config_->material_model);
  std::vector<double> mu(n, config_->mu), la(n, config_->la);
  DeviceArray<int> dt;
  DeviceArray<double> dmu, dla;
  dt.upload(t, st);
  dmu.upload(mu, st);
  dla.upload(la, st);
  mfemhip::launch_batch_material(n, dt.data(), dmu.data(), dla.data(), s, psd_fix ? 1 : 0, E, g,
                                 H, st);
  MFEM_HIP_CHECK(hipGetLastError());
  MFEM_HIP_CHECK(hipStreamSynchronize(st));
}

MaterialModelFactory::MaterialModelFactory() {
  register_type(MATERIAL_ARAP, ArapModel::name(),
                [](std::shared_ptr<MaterialConfig> c) -> std::unique_ptr<MaterialModel> {
                  return std::make_unique<ArapModel>(c);
                });
  register_type(MATERIAL_COROT, Corotational::name(),
                [](std::shared_ptr<MaterialConfig> c) -> std::unique_ptr<MaterialModel> {
                  return std::make_unique<Corotational>(c);
                });
  register_type(MATERIAL_FCR, FixedCorotational::name(),
                [](std::shared_ptr<MaterialConfig> c) -> std::unique_ptr<MaterialModel> {
                  return std::make_unique<FixedCorotational>(c);
                });
  register_type(MATERIAL_FUNG, Fung::name(),
                [](std::shared_ptr<MaterialConfig> c) -> std::unique_ptr<MaterialModel> {
                  return std::make_unique<Fung>(c);
                });
  register_type(MATERIAL_NH, Neohookean::name(),
                [](std::shared_ptr<MaterialConfig> c) -> std::unique_ptr<MaterialModel> {
                  return std::make_unique<Neohookean>(c);
                });
  register_type(MATERIAL_SNH, StableNeohookean::name(),
                [](std::shared_ptr<MaterialConfig> c) -> std::unique_ptr<MaterialModel> {
                  return std::make_unique<StableNeohookean>(c);
                });
}

bool material_supported_on_device(MaterialModelType t) {
  return t == MATERIAL_SNH || t == MATERIAL_NH || t == MATERIAL_ARAP || t == MATERIAL_COROT ||
         t == MATERIAL_FCR || t == MATERIAL_FUNG;
}

}  // namespace mfem
