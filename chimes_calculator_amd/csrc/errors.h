// Error exit shared by the C++ classes and the reference C ABIs.
//
// Like the reference (e.g. chimesFF/src/chimesFF.cpp:366-371), a fatal error
// prints its message and terminates the process with status 0; the message is
// also written to stderr, and the environment variable CHIMES_ERROR_EXIT
// selects a different exit status for callers that need to tell failures
// apart (test harnesses).
#pragma once
#include <cstdlib>
#include <iostream>
#include <string>

namespace chx {

[[noreturn]] inline void fatal_error(const std::string &msg)
{
    std::cout << msg << std::endl;
    std::cerr << msg << std::endl;
    const char *s = getenv("CHIMES_ERROR_EXIT");
    exit(s ? atoi(s) : 0);
}

}  // namespace chx
