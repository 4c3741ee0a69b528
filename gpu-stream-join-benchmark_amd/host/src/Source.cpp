// Source (src/sjb/source/Source.cpp:12-34).
#include "sjb/source/Source.hpp"

#include <atomic>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

Source::Source(uint64_t count, uint64_t bufferSize) : count(count), bufferSize(bufferSize) {}

void Source::start(Tuple *leftTuples, Tuple *rightTuples, uint64_t sourceThreads,
                   const std::shared_ptr<Windowing::BaseWindowing> &windowing) {
    const uint64_t nb = bufferSize ? count / bufferSize : 0;
    if (sourceThreads <= 1) {
        for (uint64_t i = 0; i < nb; i++) {
            windowing->onIncomingLeft(leftTuples + i * bufferSize, i);
            windowing->onIncomingRight(rightTuples + i * bufferSize, i);
        }
    } else {
        std::atomic<uint64_t> next{0};
        std::exception_ptr err;
        std::mutex errMtx;
        std::vector<std::thread> th;
        for (uint64_t t = 0; t < sourceThreads; t++)
            th.emplace_back([&] {
                try {
                    for (uint64_t i = next.fetch_add(1); i < nb; i = next.fetch_add(1)) {
                        windowing->onIncomingLeft(leftTuples + i * bufferSize, i);
                        windowing->onIncomingRight(rightTuples + i * bufferSize, i);
                    }
                } catch (...) {
                    std::lock_guard<std::mutex> g(errMtx);
                    if (!err) err = std::current_exception();
                    next.store(nb);
                }
            });
        for (auto &x : th) x.join();
        if (err) std::rethrow_exception(err);
    }
    windowing->onEndOfStream();
}
