#include "yaml_lite.hpp"

#include <fstream>
#include <sstream>
#include <stdexcept>

namespace yaml_lite {

static std::string trim(const std::string &s) {
    const size_t a = s.find_first_not_of(" \t\r\n");
    if (a == std::string::npos) return "";
    const size_t b = s.find_last_not_of(" \t\r\n");
    return s.substr(a, b - a + 1);
}

static std::string unquote(std::string s) {
    s = trim(s);
    if (s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\'')))
        return s.substr(1, s.size() - 2);
    return s;
}

// strips a trailing `# comment` that is not inside quotes
static std::string strip_comment(const std::string &s) {
    bool dq = false, sq = false;
    for (size_t i = 0; i < s.size(); i++) {
        if (s[i] == '"' && !sq) dq = !dq;
        else if (s[i] == '\'' && !dq) sq = !sq;
        else if (s[i] == '#' && !dq && !sq) return s.substr(0, i);
    }
    return s;
}

Doc parse(const std::string &text) {
    Doc doc;
    std::istringstream in(text);
    std::string line;
    int lineno = 0;
    while (std::getline(in, line)) {
        lineno++;
        line = trim(strip_comment(line));
        if (line.empty() || line == "---") continue;
        const size_t colon = line.find(':');
        if (colon == std::string::npos) throw std::invalid_argument("yaml line " + std::to_string(lineno) + ": no key");
        const std::string key = trim(line.substr(0, colon));
        std::string val = trim(line.substr(colon + 1));
        std::vector<std::string> items;
        if (!val.empty() && val.front() == '[') {
            if (val.back() != ']') throw std::invalid_argument("yaml line " + std::to_string(lineno) + ": unclosed list");
            val = val.substr(1, val.size() - 2);
            std::string cur;
            bool dq = false, sq = false;
            for (char ch : val) {
                if (ch == '"' && !sq) dq = !dq;
                if (ch == '\'' && !dq) sq = !sq;
                if (ch == ',' && !dq && !sq) {
                    if (!trim(cur).empty()) items.push_back(unquote(cur));
                    cur.clear();
                } else {
                    cur += ch;
                }
            }
            if (!trim(cur).empty()) items.push_back(unquote(cur));
        } else {
            items.push_back(unquote(val));
        }
        doc[key] = items;
    }
    return doc;
}

Doc load_file(const std::string &path) {
    std::ifstream f(path);
    if (!f) throw std::invalid_argument("cannot open " + path);
    std::stringstream ss;
    ss << f.rdbuf();
    return parse(ss.str());
}

}  // namespace yaml_lite
