#include "ParameterReader.h"

#include <cctype>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <sstream>

namespace {
// the reference's trim removes every space and tab (src/arsenal.cpp:627-640)
std::string strip_blanks(const std::string& s) {
    std::string out;
    out.reserve(s.size());
    for (char c : s)
        if (c != ' ' && c != '\t' && c != '\r') out.push_back(c);
    return out;
}
std::string lower(std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}
double first_number(const std::string& s) {
    std::stringstream ss(s + " ");
    double v = 0.;
    ss >> v;
    return v;
}
}  // namespace

long ParameterReader::find_(const std::string& name) const {
    const std::string key = lower(strip_blanks(name));
    for (size_t i = 0; i < names_.size(); i++)
        if (names_[i] == key) return static_cast<long>(i);
    return -1;
}

void ParameterReader::phraseOneLine(std::string str, std::string commentSymbol) {
    if (strip_blanks(str).empty()) return;
    const std::string eq = str.substr(0, str.find(commentSymbol));
    if (strip_blanks(eq).empty()) return;
    const size_t pos = eq.find('=');
    if (pos == std::string::npos) return;
    setVal(eq.substr(0, pos), first_number(strip_blanks(eq.substr(pos + 1))));
}

void ParameterReader::readFromFile(std::string filename, std::string commentSymbol) {
    std::ifstream f(filename);
    if (!f) {
        std::cout << "ParameterReader::readFromFile error: file " << filename
                  << " does not exist." << std::endl;
        std::exit(-1);
    }
    std::string line;
    while (std::getline(f, line)) phraseOneLine(line, commentSymbol);
}

void ParameterReader::readFromArguments(long argc, char* argv[], std::string commentSymbol,
                                        long start_from) {
    for (long i = start_from; i < argc; i++) phraseOneLine(argv[i], commentSymbol);
}

bool ParameterReader::exist(std::string name) { return find_(name) != -1; }

void ParameterReader::setVal(std::string name, double value) {
    const long i = find_(name);
    if (i == -1) {
        names_.push_back(lower(strip_blanks(name)));
        values_.push_back(value);
    } else {
        values_[i] = value;
    }
}

double ParameterReader::getVal(std::string name) {
    const long i = find_(name);
    if (i != -1) return values_[i];
    std::cout << "ParameterReader::getVal error: parameter with name " << name
              << " not found." << std::endl;
    std::exit(1);
}

double ParameterReader::getVal(std::string name, double defaultValue) {
    const long i = find_(name);
    if (i != -1) return values_[i];
    std::cout << "ParameterReader::getVal error: parameter with name " << name
              << " not found." << std::endl;
    std::cout << "using defaultValue: " << defaultValue << std::endl;
    return defaultValue;
}

void ParameterReader::echo() {
    for (size_t i = 0; i < names_.size(); i++) std::cout << names_[i] << "=" << values_[i] << "  ";
    if (!names_.empty()) std::cout << std::endl;
}
