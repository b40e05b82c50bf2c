// xml.h -- small XML reader for the parameter files: elements, attributes,
// character data, comments, processing instructions and CDATA; entity
// references &lt; &gt; &amp; &quot; &apos; and numeric ones.  Enough for the
// queries ParameterHandler makes (ParameterHandler.cpp:56,115,144,190,221,
// 270,289).
#ifndef ATUS_AMD_XML_H
#define ATUS_AMD_XML_H

#include <map>
#include <memory>
#include <string>
#include <vector>

namespace atus_host {

struct XmlNode {
  std::string name;
  std::vector<std::pair<std::string, std::string>> attrs;
  // children in document order; text runs are stored as nodes with an empty name
  std::vector<std::unique_ptr<XmlNode>> children;
  std::string text;  // for text nodes

  bool is_text() const { return name.empty(); }
  const std::string* attr(const std::string& key) const;
  // text of the first character-data child, "" if none (pugixml child_value)
  std::string child_value() const;
  bool has_text_child() const;
  // all element descendants in document order
  void descendants(std::vector<const XmlNode*>& out) const;
  std::vector<const XmlNode*> elements() const;
};

// parses a file; throws std::string on I/O or syntax errors
std::unique_ptr<XmlNode> xml_parse_file(const std::string& path);
std::unique_ptr<XmlNode> xml_parse_string(const std::string& text);

}  // namespace atus_host

#endif
