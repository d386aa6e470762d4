// ParmParse.hpp -- inputs-file database with the syntax of AMReX ParmParse as
// MuPhFASa's inputs files use it (src/IO/InputParser.cpp and the files under
// example_inputs/ and tests/regression/ of the reference):
//
//   prefix.key = value [value ...]   # comment
//
// Values are whitespace-separated tokens (a double-quoted token may contain
// spaces).  A key defined twice keeps the last definition.  Command-line
// arguments of the form key=value (after the inputs file) override the file.
#pragma once

#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace muphfasa {

class ParmParse {
public:
    ParmParse() = default;

    // Parse the text of an inputs file; later definitions replace earlier ones.
    void addText(const std::string& text);
    void addFile(const std::string& path);
    // "key=value" or "key = v1 v2" overrides (e.g. from argv).
    void addOverride(const std::string& assignment);
    void set(const std::string& key, const std::vector<std::string>& vals) { m_db[key] = vals; }

    bool contains(const std::string& key) const { return m_db.count(key) != 0; }
    const std::vector<std::string>* find(const std::string& key) const;

    // Typed access; query* return false when absent, get* throw when absent.
    bool query(const std::string& key, int& v) const;
    bool query(const std::string& key, double& v) const;
    bool query(const std::string& key, std::string& v) const;
    bool queryarr(const std::string& key, std::vector<int>& v) const;
    bool queryarr(const std::string& key, std::vector<double>& v) const;
    bool queryarr(const std::string& key, std::vector<std::string>& v) const;
    template <class T>
    void get(const std::string& key, T& v) const
    {
        if (!query(key, v)) throw std::runtime_error("ParmParse: required key '" + key + "' missing");
    }
    template <class T>
    void getarr(const std::string& key, std::vector<T>& v) const
    {
        if (!queryarr(key, v)) throw std::runtime_error("ParmParse: required key '" + key + "' missing");
    }

    const std::map<std::string, std::vector<std::string>>& table() const { return m_db; }

private:
    std::map<std::string, std::vector<std::string>> m_db;
};

}  // namespace muphfasa
