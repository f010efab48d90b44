// The HIP device a trainer or renderer runs on: the role of the reference's
// VkContext (engine/include/gs/vk_context.h:130-227) for this path — device
// selection, errors, synchronisation.  Buffers live in each trainer's
// workspace (gs_hip.h).
#pragma once

#include <stdexcept>
#include <string>

namespace gs {

// A failed HIP call or kernel launch (the reference's VulkanError).
class HipError : public std::runtime_error {
 public:
  HipError(int code, const std::string& message) : std::runtime_error(message), code(code) {}
  int code;
};

// Turns a gs_hip_* status into the exception category the reference throws:
// std::invalid_argument, std::logic_error, std::runtime_error, or HipError.
void check_status(int status);

class Device {
 public:
  explicit Device(int index = 0);
  const std::string& name() const { return name_; }
  int index() const { return index_; }
  void synchronize() const;

 private:
  int index_ = 0;
  std::string name_;
};

}  // namespace gs
