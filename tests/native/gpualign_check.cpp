// Drives GpuAlign (tweakseq_amd/qt) built against the stand-in Qt headers.
//   gpualign_check command FIN FOUT EXEC   -> the command line, one token a line
//   gpualign_check render                  -> per stdin line "a b cigar": "ga gb" or "FAIL"
//   gpualign_check pair A B                -> "OK ga gb" or "ERR message"
#include <cstdio>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>

#include "GpuAlign.h"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  GpuAlign tool;
  const std::string cmd = argv[1];
  if (cmd == "command" && argc == 5) {
    QString fin(argv[2]), fout(argv[3]), exec;
    QStringList args;
    tool.setExecutable(QString(argv[4]));
    tool.makeCommand(fin, fout, exec, args);
    std::cout << "name " << tool.name().str() << "\nexec " << exec.str() << "\n";
    for (const QString &a : args) std::cout << "arg " << a.str() << "\n";
    std::cout << "stdout " << (tool.usesStdOut() ? 1 : 0) << "\n";
    return 0;
  }
  if (cmd == "render") {
    std::string line;
    while (std::getline(std::cin, line)) {
      std::istringstream in(line);
      std::string a, b, c;
      in >> a >> b >> c;
      if (a == "-") a.clear();
      if (b == "-") b.clear();
      if (c == "-") c.clear();
      QString ga, gb;
      if (GpuAlign::render(QByteArray(a), QByteArray(b), c.data(), (int)c.size(), ga, gb))
        std::cout << (ga.size() ? ga.str() : "-") << " " << (gb.size() ? gb.str() : "-") << "\n";
      else
        std::cout << "FAIL\n";
    }
    return 0;
  }
  if (cmd == "pair" && argc == 4) {
    QString ga, gb, err;
    if (tool.alignPair(QString(argv[2]), QString(argv[3]), ga, gb, err))
      std::cout << "OK " << ga.str() << " " << gb.str() << "\n";
    else
      std::cout << "ERR " << err.str() << "\n";
    return 0;
  }
  return 2;
}
