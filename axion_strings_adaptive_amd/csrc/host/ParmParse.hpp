// ParmParse.hpp -- reader for the simulator's parameter files.
//
// Same text format and semantics the reference's parameter files use (AMReX
// ParmParse through GRTeclyn's GRParmParse): one `key = v1 v2 ...` per line,
// `#` starts a comment, later definitions override earlier ones, and
// command-line `key=value` arguments override the file.  get() of a missing
// key and error() throw ParamError with the key's full name, like
// GRParmParse::error aborting with the offending key.
#pragma once

#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace axs_host
{

struct ParamError : std::runtime_error
{
    using std::runtime_error::runtime_error;
};

class ParamTable
{
  public:
    void add_line(const std::string &raw)
    {
        std::string line = raw.substr(0, raw.find('#'));
        const auto eq = line.find('=');
        if (eq == std::string::npos)
            return;
        std::string key = trim(line.substr(0, eq));
        if (key.empty())
            return;
        std::vector<std::string> vals;
        std::istringstream is(line.substr(eq + 1));
        std::string tok;
        while (is >> tok)
            vals.push_back(tok);
        m_values[key] = vals;
    }

    void read_file(const std::string &path)
    {
        std::ifstream f(path);
        if (!f)
            throw ParamError("cannot open parameter file " + path);
        std::string line;
        while (std::getline(f, line))
            add_line(line);
    }

    bool contains(const std::string &key) const
    {
        return m_values.count(key) != 0;
    }

    const std::vector<std::string> &values(const std::string &key) const
    {
        auto it = m_values.find(key);
        if (it == m_values.end() || it->second.empty())
            throw ParamError(key + ": required parameter is missing");
        return it->second;
    }

    void set(const std::string &key, const std::string &value)
    {
        add_line(key + " = " + value);
    }

    const std::map<std::string, std::vector<std::string>> &all() const
    {
        return m_values;
    }

  private:
    static std::string trim(const std::string &s)
    {
        const auto b = s.find_first_not_of(" \t\r\n");
        if (b == std::string::npos)
            return "";
        const auto e = s.find_last_not_of(" \t\r\n");
        return s.substr(b, e - b + 1);
    }
    std::map<std::string, std::vector<std::string>> m_values;
};

// A prefix view of the table (the reference's GRParmParse("axion_strings")).
class ParmParse
{
  public:
    ParmParse(const ParamTable &t, std::string prefix)
        : m_t(t), m_prefix(std::move(prefix))
    {
    }

    std::string full(const std::string &k) const
    {
        return m_prefix.empty() ? k : m_prefix + "." + k;
    }
    bool contains(const std::string &k) const { return m_t.contains(full(k)); }

    template <typename T> void get(const std::string &k, T &v) const
    {
        v = convert<T>(m_t.values(full(k)).front(), full(k));
    }
    template <typename T> bool query(const std::string &k, T &v) const
    {
        if (!contains(k))
            return false;
        get(k, v);
        return true;
    }
    // queryAdd: like query, keeping the caller's default when absent
    template <typename T> bool queryAdd(const std::string &k, T &v) const
    {
        return query(k, v);
    }
    template <typename T> void getarr(const std::string &k, std::vector<T> &v) const
    {
        v.clear();
        for (const auto &s : m_t.values(full(k)))
            v.push_back(convert<T>(s, full(k)));
    }
    [[noreturn]] void error(const std::string &k, const std::string &msg) const
    {
        throw ParamError(full(k) + ": " + msg);
    }

  private:
    template <typename T>
    static T convert(const std::string &s, const std::string &key)
    {
        try
        {
            if constexpr (std::is_same_v<T, bool>)
            {
                if (s == "1" || s == "true" || s == "True")
                    return true;
                if (s == "0" || s == "false" || s == "False")
                    return false;
                throw ParamError(key + ": expected a boolean, got " + s);
            }
            else if constexpr (std::is_same_v<T, std::string>)
                return s;
            else if constexpr (std::is_integral_v<T>)
                return static_cast<T>(std::stoll(s));
            else
                return static_cast<T>(std::stod(s));
        }
        catch (const std::logic_error &)
        {
            throw ParamError(key + ": cannot parse value '" + s + "'");
        }
    }

    const ParamTable &m_t;
    std::string m_prefix;
};

} // namespace axs_host
