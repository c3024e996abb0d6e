/*
 * tomocam/logger.h -- the type behind nagopt's `Logger *` parameter
 * (declared at include/optimize.h:129 of the reference).  The device
 * solvers report nothing while they iterate, so a Logger here is only a
 * sink that callers may construct and pass; it forwards text to stdout or
 * appends it to a file and is otherwise unused by this library.
 */
#ifndef LOGGER_H
#define LOGGER_H

#include <cstdio>
#include <string>

namespace tomocam {

    enum class LogMode { SILENT, STDOUT, FILE };

    class Logger {
        LogMode mode_;
        std::string path_;

      public:
        explicit Logger(LogMode mode = LogMode::STDOUT, std::string path = {}) : mode_(mode), path_(std::move(path)) {}
        LogMode get_mode() const { return mode_; }
        void log(const std::string &text) const {
            if (mode_ == LogMode::STDOUT) {
                std::fputs(text.c_str(), stdout);
            } else if (mode_ == LogMode::FILE) {
                if (std::FILE *f = std::fopen(path_.c_str(), "a")) {
                    std::fputs(text.c_str(), f);
                    std::fclose(f);
                }
            }
        }
    };

}  // namespace tomocam

#endif  // LOGGER_H
