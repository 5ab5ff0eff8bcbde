// codestring.h -- solver code strings and patch groups of the drop-in.
//
// CodeString (codestring.h/.cpp:20-218): a normalised, blank-separated list of
// code words naming the flux functors a patch asks for ("fx fy fz far far far
// far far far 0").  PatchGroups (patchgroups.cpp:27-112): patches of one type
// whose code strings agree, where a word "0" agrees with any word.  The
// iterator feeder (iteratorfeeder.cpp:29-62) hands each group to the iterators
// whose code string equals the group's.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace drnum {

class CodeString : public std::string {
 public:
  CodeString() = default;
  // CodeString(const string&) / buildFrom (codestring.cpp:33-103):
  //   '\n', ',' and ';' act as blanks; a run of blanks becomes one blank;
  //   one leading and one trailing blank are removed; leading zeros of a word
  //   are removed, a word of zeros keeps one ("  aa, 034  00\n 70;5 " gives
  //   "aa 34 0 70 5").  The reference's loop is undefined on an empty result;
  //   here it stays empty.
  CodeString(const std::string& text) { buildFrom(text); }
  CodeString(const char* text) { buildFrom(text ? std::string(text) : std::string()); }
  void buildFrom(const std::string& text);

  // compareCodeString (codestring.cpp:106-174), `this` against `other`:
  //   allsame     the strings are equal;
  //   overruled   a word of `this` is "0" where `other` has another word;
  //   underruled  a word of `other` is "0" where `this` has another word.
  // Returns the word-wise merge (the non-"0" word of each pair), or "" with all
  // three flags false when the word counts differ or two words differ
  // without either being "0".
  CodeString compareCodeString(const CodeString& other, bool& allsame, bool& overruled, bool& underruled) const;

  // countNumCodeWords / accessCodeWord (codestring.cpp:177-212): the words
  // are the pieces between single blanks, so "" holds one (empty) word
  size_t countNumCodeWords() const;
  std::string accessCodeWord(size_t i_word, bool& success) const;

 private:
  std::vector<std::string> words() const;
};

// SinglePatchGroup (patchgroups.h:38-52): patches are grid indices here
struct SinglePatchGroup {
  size_t patch_type = 0;
  std::vector<size_t> patches;
  CodeString solver_codes;
  bool solver_exceptions = false;
};

// PatchGroups::insertPatch (patchgroups.cpp:35-101): the first group of the same
// type whose codes agree takes the patch (exactly equal, or, when exceptions
// are allowed, equal up to "0" words); a "0" word of the group that the patch
// fills replaces the group's codes by the merge.  Otherwise a new group.
class PatchGroups {
 public:
  explicit PatchGroups(bool allow_solver_exceptions = true) : allow_(allow_solver_exceptions) {}
  // returns (group index, index of the patch within the group)
  std::pair<size_t, size_t> insertPatch(size_t patch, size_t patch_type, const CodeString& codes);
  std::pair<size_t, size_t> insertPatch(size_t patch, size_t patch_type, const CodeString& codes,
                                        bool allow_solver_exceptions);
  size_t accessNumPatchGroups() const { return groups_.size(); }
  const SinglePatchGroup& accessSinglePatchGroup(size_t i) const { return groups_[i]; }
  void clear() { groups_.clear(); }

 private:
  bool allow_;
  std::vector<SinglePatchGroup> groups_;
};

}  // namespace drnum
