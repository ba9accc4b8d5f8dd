// Log lines in the reference's format (App/Lib/Logger.cpp:36-60):
// "<color>[<seconds since start>][<LEVEL>] message<reset>", so the
// reference's experiment scripts can parse the output of this harness.
#pragma once
#include <stdarg.h>
#include <stdio.h>
#include <time.h>

enum TbLevel { TB_INFO = 0, TB_WARN, TB_ERROR, TB_ENCLAVE, TB_DBG };

inline double teebench_seconds_since_start() {
    static struct timespec t0 = [] {
        struct timespec t;
        clock_gettime(CLOCK_MONOTONIC, &t);
        return t;
    }();
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (t.tv_sec - t0.tv_sec) + 1e-9 * (t.tv_nsec - t0.tv_nsec);
}

inline void teebench_log(TbLevel level, const char* fmt, ...) {
    static const char* names[] = {"INFO", "WARN", "ERROR", "ENCL", "DEBUG"};
    static const char* colors[] = {"\x1b[32m", "\x1b[33m", "\x1b[31m", "\x1b[35m", "\x1b[37m"};
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    printf("%s[%8.4f][%5s] %s\x1b[0m\n", colors[level], teebench_seconds_since_start(),
           names[level], buf);
}
