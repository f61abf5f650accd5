// CSV logging of the C++ host API (drop-in for the reference's
// include/logging.hpp and src/Utilities/Logging.cpp).
#ifndef LOGGING_HPP
#define LOGGING_HPP

#include <cstdio>
#include <string>

// Opens `filename` for appending and writes `header` unless the file's first
// line already is the header (Logging.cpp:3-30).  Exits on failure.
FILE* createLogCsv(std::string filename, std::string header);
void writeLineToFile(FILE* logfile, char* dataLine);
// Host name (the reference returns a pointer to a stack buffer,
// Logging.cpp:37-48; this one returns a static copy).
const char* getMachineName();

// Column header of the 2.5D Cannon logs (src/25DCannon.cpp:407).
extern const char* const kCannonCsvHeader;

#endif
