// knp/core/uid.h - 16-byte entity identifiers (interface of knp/core-library/include/knp/core/uid.h:82-242,
// without Boost.UUID: random version-4 bytes from a process-wide generator).
#pragma once
#include <array>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <random>
#include <string>

namespace knp::core {

class UID {
public:
    // A new random UID.
    UID() { generate(); }
    // UID(false) is the nil UID (used for "no presynaptic population"); UID(true) a random one.
    explicit UID(bool random) {
        if (random) generate();
        else bytes_.fill(0);
    }
    explicit operator bool() const {
        for (auto b : bytes_)
            if (b) return true;
        return false;
    }
    bool operator==(const UID &o) const { return bytes_ == o.bytes_; }
    bool operator!=(const UID &o) const { return !(*this == o); }
    bool operator<(const UID &o) const { return bytes_ < o.bytes_; }
    explicit operator std::string() const {
        static const char *hex = "0123456789abcdef";
        std::string s;
        for (size_t i = 0; i < bytes_.size(); ++i) {
            if (i == 4 || i == 6 || i == 8 || i == 10) s.push_back('-');
            s.push_back(hex[bytes_[i] >> 4]);
            s.push_back(hex[bytes_[i] & 15]);
        }
        return s;
    }
    const std::array<uint8_t, 16> &bytes() const { return bytes_; }

private:
    void generate() {
        static std::mutex m;
        static std::mt19937_64 gen{std::random_device{}()};
        std::lock_guard<std::mutex> lock(m);
        uint64_t a = gen(), b = gen();
        std::memcpy(bytes_.data(), &a, 8);
        std::memcpy(bytes_.data() + 8, &b, 8);
        bytes_[6] = (bytes_[6] & 0x0F) | 0x40;
        bytes_[8] = (bytes_[8] & 0x3F) | 0x80;
    }
    std::array<uint8_t, 16> bytes_{};
};

struct uid_hash {
    size_t operator()(const UID &u) const {
        uint64_t a, b;
        std::memcpy(&a, u.bytes().data(), 8);
        std::memcpy(&b, u.bytes().data() + 8, 8);
        return std::hash<uint64_t>{}(a ^ (b * 0x9E3779B97F4A7C15ULL));
    }
};

using Step = uint64_t;

}  // namespace knp::core
