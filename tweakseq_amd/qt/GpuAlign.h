// GpuAlign: an alignment tool for tweakseq over the tweakalign C ABI.
//
// It subclasses tweakseq's AlignmentTool (tweakseq/Core/AlignmentTool.h:36-71)
// so that settings, menus and the process path keep working:
//  * process path (makeCommand): the tweakalign CLI with MUSCLE's argument
//    list, "-in FILE -out FILE" (tweakseq/Core/Muscle.cpp:49-53);
//  * in-process path (alignPair): a global alignment of two residue strings
//    on the GPU through ta_align_pair, rendered as two gapped rows -- what the
//    editor's realign-selection splices back (tweakseq/Core/Project.cpp:964-1026).
// Built into tweakseq next to Muscle.cpp and registered by name where the
// tools are listed (tweakseq/Core/Application.cpp:196-204,
// tweakseq/Core/Project.cpp:1142-1152, tweakseq/UI/SetupWizard.cpp:207-209).
#ifndef GPU_ALIGN_H
#define GPU_ALIGN_H

#include <QByteArray>
#include <QString>
#include <QStringList>

#include "AlignmentTool.h"

struct ta_ctx;

class GpuAlign : public AlignmentTool
{
	public:
		GpuAlign();
		virtual ~GpuAlign();

		virtual void makeCommand(QString &fin, QString &fout, QString &exec, QStringList &args);
		virtual void writeSettings(QDomDocument &doc, QDomElement &parent);
		virtual void readSettings(QDomDocument &doc);

		// Global alignment of residue strings a and b (gaps already removed)
		// on the GPU; false with err set when no GPU or the call fails.
		bool alignPair(const QString &a, const QString &b, QString &gappedA, QString &gappedB, QString &err);

		// A CIGAR over a (query: 'I' consumes it) and b ('D' consumes it) as
		// two gapped rows; false when the runs do not consume both exactly.
		static bool render(const QByteArray &a, const QByteArray &b, const char *cigar, int len,
		                   QString &gappedA, QString &gappedB);

	private:
		ta_ctx *ctx_;
};

#endif
