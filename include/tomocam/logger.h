/*
 * tomocam/logger.h -- the progress logger that nagopt accepts
 * (include/logger.h:28-72 of the reference).  The device solvers do not
 * print; a Logger passed to them is accepted and left unused.
 */
#ifndef LOGGER_H
#define LOGGER_H

#include <fstream>
#include <iostream>
#include <memory>
#include <string>

namespace tomocam {

    enum class LogMode { SILENT, STDOUT, FILE };

    class Logger {
        LogMode mode_;
        std::unique_ptr<std::ofstream> out_;

      public:
        explicit Logger(LogMode mode = LogMode::STDOUT, const std::string &filename = "") : mode_(mode) {
            if (mode_ == LogMode::FILE) {
                out_ = std::make_unique<std::ofstream>(filename);
                if (!out_->is_open()) {
                    std::cerr << "Warning: Unable to open log file '" << filename << "'. Falling back to STDOUT.\n";
                    mode_ = LogMode::STDOUT;
                }
            }
        }
        void log(const std::string &msg) {
            if (mode_ == LogMode::STDOUT) std::cout << msg;
            else if (mode_ == LogMode::FILE && out_) *out_ << msg << std::flush;
        }
        LogMode get_mode() const { return mode_; }
    };

}  // namespace tomocam

#endif  // LOGGER_H
