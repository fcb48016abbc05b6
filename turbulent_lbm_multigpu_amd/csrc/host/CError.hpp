// CError.hpp -- error accumulator with the reference's convention
// (src/lib/CError.hpp:42-80): operator() tells whether an error was recorded,
// getString() returns the accumulated text, operator<< appends.
#pragma once

#include <sstream>
#include <string>

class CError {
    std::ostringstream text_;
    bool error_ = false;

public:
    static constexpr const char *endl = "\n";

    bool operator()() const { return error_; }
    std::string getString() const { return text_.str(); }
    void clear() {
        text_.str("");
        error_ = false;
    }
    template <typename V>
    CError &operator<<(const V &v) {
        error_ = true;
        text_ << v;
        return *this;
    }
    CError &operator<<(std::ostream &(*manip)(std::ostream &)) {
        error_ = true;
        text_ << manip;
        return *this;
    }
};
