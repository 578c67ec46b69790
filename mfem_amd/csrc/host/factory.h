// Name <-> enum <-> creator registries with the reference's API
// (src/factories/factory.h:10-69): create(type, args...), names(),
// type_by_name(name), name_by_type(type), register_type(...). As in the
// reference, an unknown name maps to enum 0 with a message on stderr and an
// unknown type creates nullptr.
#pragma once
#include <iostream>
#include <map>
#include <memory>
#include <string>
#include <vector>

namespace mfem {

template <typename TypeEnum, typename T, typename... Ts>
class Factory {
 public:
  using TypeCreator = std::unique_ptr<T> (*)(Ts... args);
  virtual ~Factory() = default;

  std::unique_ptr<T> create(TypeEnum type, Ts... args) {
    auto it = creators_.find(type);
    if (it != creators_.end()) return it->second(std::move(args)...);
    std::cerr << "Factory create: type not found" << std::endl;
    return nullptr;
  }
  const std::vector<std::string>& names() const { return names_; }
  bool has_name(const std::string& name) const { return by_name_.count(name) > 0; }
  TypeEnum type_by_name(const std::string& name) const {
    auto it = by_name_.find(name);
    if (it != by_name_.end()) return it->second;
    std::cerr << "Factory type_by_name: type not found" << std::endl;
    return (TypeEnum)0;
  }
  std::string name_by_type(TypeEnum type) const {
    auto it = by_type_.find(type);
    if (it != by_type_.end()) return it->second;
    std::cerr << "Factory name_by_type: name not found" << std::endl;
    return "";
  }

 protected:
  void register_type(TypeEnum type, const std::string& name, TypeCreator func) {
    creators_.emplace(type, func);
    by_name_.emplace(name, type);
    by_type_.emplace(type, name);
    names_.push_back(name);
  }

 private:
  std::map<TypeEnum, TypeCreator> creators_;
  std::map<std::string, TypeEnum> by_name_;
  std::map<TypeEnum, std::string> by_type_;
  std::vector<std::string> names_;
};

}  // namespace mfem
