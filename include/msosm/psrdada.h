// psrdada.h -- drop-in facade for the reference's PSRDADA file reader
// (psrdada/include/psrdada.h, psrdada/psrdada.cpp): same class name,
// constructor, methods and public members.
//
// File layout read here (psrdada.cpp:18-56,87-94): a 4096-byte ASCII header of
// "KEY value" lines (NUL padded; BW and FREQ in MHz, FREQ the band centre),
// then the two polarisations' complex uint16 samples in alternating frames of
// 2048 samples: pol1 frame, pol2 frame, pol1 frame, ...
//
// Differences a caller can see: pol1_in / pol2_in are page-locked (host to
// device copies of each block run at full PCIe speed; the reference used
// plain new[]), and are released in the destructor.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <utility>

#include "globals.h"

constexpr int blockSize = 4096;    // header bytes (psrdada.h:11)
constexpr int frameSamples = 2048; // samples per polarisation frame (psrdada.h:12)

class PSRDADA {
public:
    explicit PSRDADA(std::string filename) : infileName(filename)
    {
        infile.open(filename, std::ios::binary | std::ios::ate);
        if (!infile.is_open()) {
            std::cout << "Cannot open file: " << filename << std::endl;
            std::exit(1);
        }
        filesize = static_cast<unsigned long>(infile.tellg());
        infile.seekg(0, std::ios::beg);
        outfileName = filename.substr(0, filename.find_last_of('.')) + "_dedisp_folded.txt";
        infile.read(headerBuffer, blockSize);
        if (infile.gcount() != blockSize) {
            std::cout << "File shorter than the " << blockSize << "-byte header: " << filename << std::endl;
            std::exit(1);
        }
        parse_header();
        std::cout << bw << " " << fc << " " << nbit << " " << nchan << " " << ndim << " " << npol << std::endl;
    }
    ~PSRDADA()
    {
        msosm_host_free(pol1_in);
        msosm_host_free(pol2_in);
        infile.close();
    }
    PSRDADA(const PSRDADA &) = delete;
    PSRDADA &operator=(const PSRDADA &) = delete;

    // psrdada.cpp:64-85: per-polarisation buffers of n samples and the number
    // of whole (pol1 + pol2) reads the data section holds
    void initBuffer(size_t n)
    {
        if (n % frameSamples != 0) {
            std::cout << "n must be a multiple of frameSamples (" << frameSamples << ")" << std::endl;
            std::exit(1);
        }
        msosm_host_free(pol1_in);
        msosm_host_free(pol2_in);
        void *a = nullptr, *b = nullptr;
        msosm_facade::check(msosm_host_alloc(n * sizeof(uint16_pair), &a), "PSRDADA::initBuffer");
        msosm_facade::check(msosm_host_alloc(n * sizeof(uint16_pair), &b), "PSRDADA::initBuffer");
        pol1_in = static_cast<uint16_pair *>(a);
        pol2_in = static_cast<uint16_pair *>(b);
        frameCount = static_cast<int>(n / frameSamples);
        readCount = static_cast<unsigned long>((filesize - blockSize) / (2 * n * sizeof(uint16_pair)));
        if (verbose) {
            std::cout << "readSize: " << n << std::endl;
            std::cout << "frameCount: " << frameCount << std::endl;
            std::cout << "Filesize: " << filesize << " bytes" << std::endl;
            std::cout << "readCount: " << readCount << std::endl;
        }
    }

    // psrdada.cpp:87-94: the next frameCount frame pairs, split by polarisation
    void readSamples()
    {
        const std::streamsize fb = frameSamples * static_cast<std::streamsize>(sizeof(uint16_pair));
        for (int i = 0; i < frameCount; i++) {
            infile.read(reinterpret_cast<char *>(pol1_in + static_cast<size_t>(i) * frameSamples), fb);
            infile.read(reinterpret_cast<char *>(pol2_in + static_cast<size_t>(i) * frameSamples), fb);
        }
    }

public:
    char headerBuffer[blockSize] = {};
    std::ifstream infile;
    std::string infileName;
    unsigned long filesize = 0;
    unsigned long readCount = 0;
    std::string outfileName;
    float bw = 0;  // MHz
    float fc = 0;  // MHz, band centre
    int nbit = 0;
    int nchan = 0;
    int ndim = 0;
    int npol = 0;
    int hdr_size = 0;
    uint16_pair *pol1_in = nullptr;
    uint16_pair *pol2_in = nullptr;
    int frameCount = 0;

private:
    // psrdada.cpp:30-56, 97-116: text up to the first NUL, one "KEY value"
    // pair per line, the value after the run of spaces following the key
    void parse_header()
    {
        std::string text(headerBuffer, strnlen(headerBuffer, blockSize));
        std::istringstream lines(text);
        std::string line;
        while (std::getline(lines, line)) {
            const size_t sp = line.find(' ');
            if (sp == std::string::npos)
                continue;
            const size_t vs = line.find_first_not_of(' ', sp);
            if (vs == std::string::npos)
                continue;
            const std::string key = line.substr(0, sp), val = line.substr(vs);
            if (verbose)
                std::cout << key << " " << val << std::endl;
            if (key == "BW")
                bw = std::stof(val);
            else if (key == "FREQ")
                fc = std::stof(val);
            else if (key == "NBIT")
                nbit = std::stoi(val);
            else if (key == "NCHAN")
                nchan = std::stoi(val);
            else if (key == "NDIM")
                ndim = std::stoi(val);
            else if (key == "NPOL")
                npol = std::stoi(val);
        }
    }
};
