// Logging and fatal errors with the reference's conventions
// (src/error_handler.{hpp,cpp}): fatal() logs and throws std::runtime_error,
// warning() prints to stderr, log() appends a timestamped line to .smax_log.
#pragma once

#include <fstream>
#include <stdexcept>
#include <string>

namespace SMAX {

class ErrorHandler {
  public:
    static void initialize_log(const std::string &filename);
    static void close_log();
    [[noreturn]] static void fatal(const std::string &message);
    static void warning(const std::string &message);
    static void log(const std::string &message);
    [[noreturn]] static void kernel_dne(const std::string &kernel_name);
};

} // namespace SMAX
