// GpuAlign over the tweakalign C ABI (see GpuAlign.h).
#include "GpuAlign.h"

#include "tweakalign.h"

GpuAlign::GpuAlign() : ctx_(NULL)
{
	name_ = "tweakalign";
	version_ = "1.0.0";
	usesStdOut_ = false;
	preferred_ = false;
}

GpuAlign::~GpuAlign()
{
	if (ctx_) ta_destroy(ctx_);
}

void GpuAlign::makeCommand(QString &fin, QString &fout, QString &exec, QStringList &args)
{
	exec = executable_;
	args << "-in" << fin << "-out" << fout;
}

void GpuAlign::writeSettings(QDomDocument &doc, QDomElement &parent)
{
	AlignmentTool::writeSettings(doc, parent);
}

void GpuAlign::readSettings(QDomDocument &doc)
{
	AlignmentTool::readSettings(doc);
}

bool GpuAlign::render(const QByteArray &a, const QByteArray &b, const char *cigar, int len, QString &ga,
                      QString &gb)
{
	ga.clear();
	gb.clear();
	int i = 0, j = 0;
	long run = 0;
	for (int k = 0; k < len; ++k) {
		const char c = cigar[k];
		if (c >= '0' && c <= '9') {
			run = 10 * run + (c - '0');
			if (run > a.size() + b.size()) return false;
			continue;
		}
		if ((c != 'M' && c != 'I' && c != 'D') || run == 0) return false;
		for (long t = 0; t < run; ++t) {
			if (c != 'D' && i >= a.size()) return false;
			if (c != 'I' && j >= b.size()) return false;
			ga += (c == 'D') ? QChar('-') : QChar(a[i++]);
			gb += (c == 'I') ? QChar('-') : QChar(b[j++]);
		}
		run = 0;
	}
	return run == 0 && i == a.size() && j == b.size();
}

bool GpuAlign::alignPair(const QString &a, const QString &b, QString &ga, QString &gb, QString &err)
{
	if (!ctx_) {
		const int rc = ta_create(&ctx_, 0);
		if (rc != TA_OK) {
			err = ta_strerror(rc);
			ctx_ = NULL;
			return false;
		}
	}
	const QByteArray A = a.toLatin1(), B = b.toLatin1();
	const ta_params p = {TA_GLOBAL, TA_PROTEIN, -10, -1, 0, 0, 1, TA_LANES_AUTO};
	ta_result r;
	int64_t used = 0;
	QByteArray cig(4 * (A.size() + B.size()) + 16, '\0');  // one run per residue at most
	const int rc = ta_align_pair(ctx_, &p, A.constData(), A.size(), B.constData(), B.size(), &r, cig.data(),
	                             cig.size(), &used);
	if (rc != TA_OK) {
		err = ta_strerror(rc);
		return false;
	}
	if (!render(A, B, cig.constData() + r.cigar_off, r.cigar_len, ga, gb)) {
		err = "malformed CIGAR";
		return false;
	}
	return true;
}
