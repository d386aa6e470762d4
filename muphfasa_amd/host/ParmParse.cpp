#include "ParmParse.hpp"

#include <cstdlib>
#include <fstream>
#include <sstream>

namespace muphfasa {

namespace {

std::vector<std::string> tokenize(const std::string& s)
{
    std::vector<std::string> out;
    size_t i = 0;
    while (i < s.size()) {
        while (i < s.size() && (s[i] == ' ' || s[i] == '\t' || s[i] == '\r')) ++i;
        if (i >= s.size()) break;
        if (s[i] == '"') {
            size_t j = s.find('"', i + 1);
            if (j == std::string::npos) throw std::runtime_error("ParmParse: unterminated quote");
            out.push_back(s.substr(i + 1, j - i - 1));
            i = j + 1;
        } else {
            size_t j = i;
            while (j < s.size() && s[j] != ' ' && s[j] != '\t' && s[j] != '\r') ++j;
            out.push_back(s.substr(i, j - i));
            i = j;
        }
    }
    return out;
}

std::string strip_comment(const std::string& line)
{
    bool inq = false;
    for (size_t i = 0; i < line.size(); ++i) {
        if (line[i] == '"') inq = !inq;
        if (line[i] == '#' && !inq) return line.substr(0, i);
    }
    return line;
}

std::string trim(const std::string& s)
{
    size_t a = s.find_first_not_of(" \t\r");
    if (a == std::string::npos) return "";
    size_t b = s.find_last_not_of(" \t\r");
    return s.substr(a, b - a + 1);
}

}  // namespace

void ParmParse::addText(const std::string& text)
{
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        ++lineno;
        line = strip_comment(line);
        if (trim(line).empty()) continue;
        const size_t eq = line.find('=');
        if (eq == std::string::npos)
            throw std::runtime_error("ParmParse: line " + std::to_string(lineno) +
                                     " has no '=': " + trim(line));
        const std::string key = trim(line.substr(0, eq));
        if (key.empty()) throw std::runtime_error("ParmParse: empty key on line " + std::to_string(lineno));
        m_db[key] = tokenize(line.substr(eq + 1));
    }
}

void ParmParse::addFile(const std::string& path)
{
    std::ifstream in(path);
    if (!in) throw std::runtime_error("ParmParse: cannot open inputs file '" + path + "'");
    std::stringstream ss;
    ss << in.rdbuf();
    addText(ss.str());
}

void ParmParse::addOverride(const std::string& assignment) { addText(assignment + "\n"); }

const std::vector<std::string>* ParmParse::find(const std::string& key) const
{
    auto it = m_db.find(key);
    return it == m_db.end() ? nullptr : &it->second;
}

static double to_double(const std::string& key, const std::string& s)
{
    char* end = nullptr;
    const double v = std::strtod(s.c_str(), &end);
    if (!end || *end != '\0') throw std::runtime_error("ParmParse: '" + key + "': not a number: " + s);
    return v;
}

static int to_int(const std::string& key, const std::string& s)
{
    char* end = nullptr;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (!end || *end != '\0') throw std::runtime_error("ParmParse: '" + key + "': not an integer: " + s);
    return (int)v;
}

bool ParmParse::query(const std::string& key, int& v) const
{
    auto p = find(key);
    if (!p || p->empty()) return false;
    v = to_int(key, (*p)[0]);
    return true;
}

bool ParmParse::query(const std::string& key, double& v) const
{
    auto p = find(key);
    if (!p || p->empty()) return false;
    v = to_double(key, (*p)[0]);
    return true;
}

bool ParmParse::query(const std::string& key, std::string& v) const
{
    auto p = find(key);
    if (!p || p->empty()) return false;
    v = (*p)[0];
    return true;
}

bool ParmParse::queryarr(const std::string& key, std::vector<int>& v) const
{
    auto p = find(key);
    if (!p) return false;
    v.clear();
    for (auto& s : *p) v.push_back(to_int(key, s));
    return true;
}

bool ParmParse::queryarr(const std::string& key, std::vector<double>& v) const
{
    auto p = find(key);
    if (!p) return false;
    v.clear();
    for (auto& s : *p) v.push_back(to_double(key, s));
    return true;
}

bool ParmParse::queryarr(const std::string& key, std::vector<std::string>& v) const
{
    auto p = find(key);
    if (!p) return false;
    v = *p;
    return true;
}

}  // namespace muphfasa
