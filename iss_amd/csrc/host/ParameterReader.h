// Flat "name = value  # comment" parameter store with the reference's
// semantics (src/ParameterReader.h:14-57, src/ParameterReader.cpp:23-187):
// names are lower-cased, spaces and tabs are removed from names and values,
// later assignments overwrite earlier ones, getVal(name) exits when the name
// is missing and getVal(name, default) falls back to the default.
#ifndef ISS_AMD_PARAMETERREADER_H_
#define ISS_AMD_PARAMETERREADER_H_

#include <string>
#include <vector>

class ParameterReader {
 private:
    std::vector<std::string> names_;
    std::vector<double> values_;
    long find_(const std::string& name) const;

 public:
    ParameterReader() = default;
    ~ParameterReader() = default;

    void phraseOneLine(std::string str, std::string commentSymbol = "#");
    void readFromFile(std::string filename, std::string commentSymbol = "#");
    void readFromArguments(long argc, char* argv[], std::string commentSymbol = "#",
                           long start_from = 1);
    bool exist(std::string name);
    void setVal(std::string name, double value);
    double getVal(std::string name);
    double getVal(std::string name, double defaultValue);
    void echo();
};

#endif  // ISS_AMD_PARAMETERREADER_H_
