// The interface of tweakseq's AlignmentTool (tweakseq/Core/AlignmentTool.h:36-71),
// declared again so that GpuAlign, which subclasses it, compiles in a CPU
// test; member names and signatures must be the host's.  The bodies are
// trivial here (the host's AlignmentTool.cpp has the real ones).
#ifndef FAKE_ALIGNMENT_TOOL_H
#define FAKE_ALIGNMENT_TOOL_H
#include <QString>
#include <QStringList>
#include <QDomDocument>

class AlignmentTool {
 public:
  AlignmentTool() : preferred_(false), usesStdOut_(false) {}
  virtual ~AlignmentTool() {}
  QString name() { return name_; }
  QString version() { return version_; }
  QString executable() { return executable_; }
  void setExecutable(QString e) { executable_ = e; }
  void setPreferred(bool pref) { preferred_ = pref; }
  bool preferred() { return preferred_; }
  bool usesStdOut() { return usesStdOut_; }
  virtual void makeCommand(QString &, QString &, QString &, QStringList &) {}
  virtual void writeSettings(QDomDocument &, QDomElement &) {}
  virtual void readSettings(QDomDocument &) {}
 protected:
  QString name_;
  QString version_;
  QString executable_;
  bool preferred_;
  bool usesStdOut_;
};
#endif
