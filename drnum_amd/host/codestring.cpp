// codestring.cpp -- CodeString and PatchGroups (codestring.cpp:33-212,
// patchgroups.cpp:27-112 of the reference).
#include "codestring.h"

namespace drnum {

void CodeString::buildFrom(const std::string& text) {
  std::string s;
  s.reserve(text.size());
  bool after_blank = false;
  for (char c : text) {
    if (c == '\n' || c == ',' || c == ';') c = ' ';
    if (c == ' ') {
      if (after_blank) continue;  // one blank per run
      after_blank = true;
    } else {
      after_blank = false;
    }
    s.push_back(c);
  }
  if (!s.empty() && s.front() == ' ') s.erase(0, 1);
  if (!s.empty() && s.back() == ' ') s.pop_back();
  // drop the zeros that start a word and are followed by another character of it
  std::string out;
  out.reserve(s.size());
  bool word_start = true;
  for (size_t i = 0; i < s.size(); ++i) {
    const char c = s[i];
    if (c == ' ') {
      word_start = true;
      out.push_back(c);
      continue;
    }
    const bool last_of_word = (i + 1 == s.size()) || s[i + 1] == ' ';
    if (word_start && c == '0' && !last_of_word) continue;
    word_start = false;
    out.push_back(c);
  }
  assign(out);
}

std::vector<std::string> CodeString::words() const {
  std::vector<std::string> w(1);
  for (char c : *this) {
    if (c == ' ')
      w.emplace_back();
    else
      w.back().push_back(c);
  }
  return w;
}

CodeString CodeString::compareCodeString(const CodeString& other, bool& allsame, bool& overruled,
                                         bool& underruled) const {
  allsame = overruled = underruled = false;
  if (compare(other) == 0) {
    allsame = true;
    return other;
  }
  const std::vector<std::string> a = words(), b = other.words();
  CodeString merged;
  bool differ = a.size() != b.size();
  for (size_t i = 0; i < a.size() && !differ; ++i) {
    std::string m;
    if (a[i] == b[i]) {
      m = a[i];
    } else if (a[i] == "0") {
      overruled = true;
      m = b[i];
    } else if (b[i] == "0") {
      underruled = true;
      m = a[i];
    } else {
      differ = true;
      break;
    }
    if (i > 0) merged.push_back(' ');
    merged.append(m);
  }
  if (differ) {
    overruled = underruled = false;
    return CodeString();
  }
  return merged;
}

size_t CodeString::countNumCodeWords() const { return words().size(); }

std::string CodeString::accessCodeWord(size_t i_word, bool& success) const {
  const std::vector<std::string> w = words();
  success = i_word < w.size();
  return success ? w[i_word] : std::string();
}

std::pair<size_t, size_t> PatchGroups::insertPatch(size_t patch, size_t patch_type, const CodeString& codes) {
  return insertPatch(patch, patch_type, codes, allow_);
}

std::pair<size_t, size_t> PatchGroups::insertPatch(size_t patch, size_t patch_type, const CodeString& codes,
                                                   bool allow_solver_exceptions) {
  for (size_t g = 0; g < groups_.size(); ++g) {
    SinglePatchGroup& G = groups_[g];
    if (G.patch_type != patch_type) continue;
    bool same, over, under;
    const CodeString merged = codes.compareCodeString(G.solver_codes, same, over, under);
    if (!(same || ((over || under) && allow_solver_exceptions))) continue;
    G.patches.push_back(patch);
    if (under) G.solver_codes = merged;
    if (under || over) G.solver_exceptions = true;
    return {g, G.patches.size() - 1};
  }
  SinglePatchGroup G;
  G.patch_type = patch_type;
  G.solver_codes = codes;
  G.patches.push_back(patch);
  groups_.push_back(std::move(G));
  return {groups_.size() - 1, 0};
}

}  // namespace drnum
