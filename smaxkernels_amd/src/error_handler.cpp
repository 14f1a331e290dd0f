#include "error_handler.hpp"

#include <chrono>
#include <ctime>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <sstream>

namespace SMAX {

namespace {
std::ofstream g_log;
std::mutex g_mutex;
long g_lines = 0;
constexpr long kLineWarning = 100000;   // error_handler.cpp:12
constexpr long kLineLimit = 1000000;    // error_handler.cpp:13

std::string now_stamp() {
    auto now = std::chrono::system_clock::now();
    std::time_t t = std::chrono::system_clock::to_time_t(now);
    std::tm tm{};
    localtime_r(&t, &tm);
    auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                  now.time_since_epoch()) % 1000000;
    std::ostringstream o;
    o << "[" << std::put_time(&tm, "%Y-%m-%d %H:%M:%S") << "." << std::setw(6)
      << std::setfill('0') << us.count() << "] ";
    return o.str();
}
} // namespace

void ErrorHandler::initialize_log(const std::string &filename) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_log.is_open()) return;
    g_log.open(filename, std::ios::out);
    if (!g_log.is_open())
        throw std::runtime_error("Failed to open log file: " + filename);
}

void ErrorHandler::close_log() {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (g_log.is_open()) g_log.close();
}

void ErrorHandler::log(const std::string &message) {
    std::lock_guard<std::mutex> lk(g_mutex);
    if (!g_log.is_open()) return;
    g_log << now_stamp() << message << '\n';
    ++g_lines;
    if (g_lines % kLineWarning == 0)
        std::cerr << "Warning: " << kLineWarning << " lines in log file.\n";
    if (g_lines >= kLineLimit) g_log.close();
}

void ErrorHandler::fatal(const std::string &message) {
    const std::string full = "[FATAL] " + message;
    log(full);
    throw std::runtime_error(full);
}

void ErrorHandler::warning(const std::string &message) {
    const std::string full = "[WARNING] " + message;
    std::cerr << full << std::endl;
    log(full);
}

void ErrorHandler::kernel_dne(const std::string &kernel_name) {
    fatal("Kernel: '" + kernel_name + "' does not exist.");
}

} // namespace SMAX
