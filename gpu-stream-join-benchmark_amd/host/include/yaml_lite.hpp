// Minimal reader for the experiment YAML files of the reference driver
// (configs/*.yaml: one `key: value` per line, value a flow list
// `[ "a", "b" ]` / `[ 1, 2 ]` or a scalar, `#` comments).  yaml-cpp is not
// available here; this covers exactly that subset.
#pragma once
#include <map>
#include <string>
#include <vector>

namespace yaml_lite {
using Doc = std::map<std::string, std::vector<std::string>>;
// Throws std::invalid_argument on malformed lines.
Doc parse(const std::string &text);
Doc load_file(const std::string &path);
}  // namespace yaml_lite
